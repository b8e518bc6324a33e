"""The reference's time loop over a Solver, and its input file.

``Run`` mirrors ``run`` (multideal_gpu.F90:1659-1851) for one rank: the dt recompute every
``iter_dt_recompute`` steps (eval_aux of the interior, then compute_dt), ``advance_solution``,
eval_aux of the interior, every ``print_control`` steps the monitors of ``compute_residual``
(:1586-1618: the momentum residual sqrt(sum (fln_u/dt)^2 / cells), the NaN search when it is
NaN, the extrema of rho, T and p) and the progress line of ``print_progress``
(:1933-1972), the time update and the validation writers of ``manage_output`` (:2465-2564):

* every case: ``output_1d.dat`` (x and every auxiliary variable along the first line,
  ghosts included) whenever the time since the last write reaches ``dtsave``;
* the 0-D reactor (flow_init 8): ``output_batch.dat`` rows (time, rho, T) every 10 steps;
* the reactive shock tube (flow_init 9): ``output_1d_{170,190,230}micro_s.dat`` (x, T, u and
  the third auxiliary variable) at the first step reaching each time.

A NaN residual stops the run with the location of the first bad cell, as the reference
aborts with 'BOOM'.  ``read_ini`` reads the reference's ``multideal.ini``; ``Run.from_ini``
builds the mechanism tables, grid, numerics and initial state of the supported flow_init
cases (8: init_reactor, multideal.F90:5608-5669; 9: init_reactive_tube, :5305-5381) from it.
"""
from __future__ import annotations

import dataclasses
import math
import pathlib
import time as _time
from typing import Dict, List, Optional

import numpy as np

from . import cases, solver as solvermod
from .mechanism import Mechanism, Tables

FMT_1D = "{:30.15E}"


def read_ini(path) -> Dict[str, Dict[str, str]]:
    """The sections of a STREAmS ini file: ``key = value`` pairs, ';' starts a comment."""
    out: Dict[str, Dict[str, str]] = {}
    sec = None
    for raw in pathlib.Path(path).read_text().splitlines():
        line = raw.split(";", 1)[0].strip()
        if not line:
            continue
        if line.startswith("[") and line.endswith("]"):
            sec = line[1:-1].strip().lower()
            out.setdefault(sec, {})
            continue
        if "=" in line and sec is not None:
            k, v = line.split("=", 1)
            out[sec][k.strip().lower()] = v.strip()
    return out


def _num(d, k, default=None, kind=float):
    if k not in d:
        if default is None:
            raise KeyError(k)
        return default
    return kind(float(d[k])) if kind is int else kind(d[k])


@dataclasses.dataclass
class Controls:
    cfl: float = 0.5                 # > 0 CFL number, < 0 minus the fixed time step
    iter_dt_recompute: int = 1
    num_iter: int = 1
    print_control: int = 1
    dtsave: float = 1e6
    flow_init: int = -1


class BoomError(RuntimeError):
    pass


class Run:
    """One rank's time loop.  ``outdir`` receives the output files (None: no files)."""

    def __init__(self, solver: solvermod.Solver, controls: Controls, outdir: Optional[str] = None):
        self.s = solver
        self.c = controls
        self.out = pathlib.Path(outdir) if outdir is not None else None
        if self.out is not None:
            self.out.mkdir(parents=True, exist_ok=True)
        self.icyc = 0
        self.icyc0 = 0
        self.residual_rhou = 0.0
        self.extrema = dict(rhomin=0.0, rhomax=0.0, tmin=0.0, tmax=0.0, pmin=0.0, pmax=0.0)
        self.time_from_last_write = 0.0
        self.istore = 1
        self.tube_done = [0, 0, 0]
        self.progress: List[tuple] = []
        self.batch_rows: List[tuple] = []
        self.averaged_timing = 0.0

    # ------------------------------------------------------------------ monitors
    def compute_residual(self):
        """compute_residual (multideal_gpu.F90:1586-1618) for one rank."""
        s = self.s
        be = s.be
        res = np.zeros(1)
        be.call("compute_residual", s.nx, s.ny, s.nz, s.ng, s.nv, s.fln, s.dt, res, None, be.stream)
        if math.isnan(res[0]):
            n, ijk = np.zeros(1, dtype=np.int32), np.zeros(3, dtype=np.int32)
            be.call("find_boom", s.nx, s.ny, s.nz, s.ng, s.nv, s.fln, s.dt, None, 0, 0, n, ijk, be.stream)
            self.boom = (int(n[0]), tuple(int(v) for v in ijk))
        self.residual_rhou = math.sqrt(res[0] / (float(s.nx) * float(s.ny) * float(s.nz)))
        v = [np.zeros(1) for _ in range(6)]
        be.call("compute_rho_t_p_minmax", s.ns, s.nx, s.ny, s.nz, s.ng, s.nv_aux, s.w_aux, *v, None, be.stream)
        self.extrema = dict(zip(("rhomin", "rhomax", "tmin", "tmax", "pmin", "pmax"), (float(a[0]) for a in v)))

    # ------------------------------------------------------------------ loop
    def step(self):
        s, c = self.s, self.c
        t0 = _time.perf_counter()
        self.icyc += 1
        if (self.icyc - self.icyc0) % c.iter_dt_recompute == 0:
            s.compute_aux(central=1, ghost=0)
            s.compute_dt(c.cfl)
        s.advance()
        s.compute_aux(central=1, ghost=0)
        if s.num.enable_chemistry > 0 and self.time_from_last_write + s.dt >= c.dtsave:
            s.compute_chem_aux()        # the diagnostics an output will write
        if (self.icyc - self.icyc0) % c.print_control == 0:
            self.compute_residual()
            if math.isnan(self.residual_rhou):
                n, ijk = getattr(self, "boom", (0, (0, 0, 0)))
                raise BoomError(f"BOOM: NaN residual at step {self.icyc}; {n} bad cell(s), first at i,j,k = {ijk}")
        # advance() has added dt to the solver's time (the reference adds it after the
        # monitors, before manage_output)
        self.manage_output()
        s.be.sync()
        step_time = _time.perf_counter() - t0
        if (self.icyc - self.icyc0) % c.print_control == 0:
            self.print_progress(step_time)

    def run(self, num_iter: Optional[int] = None):
        """The integration loop of run; returns the averaged wall time per step."""
        s = self.s
        n = self.c.num_iter if num_iter is None else num_iter
        self.update_start()
        t0 = _time.perf_counter()
        for _ in range(n):
            self.step()
        s.be.sync()
        self.averaged_timing = (_time.perf_counter() - t0) / max(1, n)
        return self.averaged_timing

    def update_start(self):
        """The set-up of run before the loop: physical ghosts, auxiliaries with the sensor
        exchange, and the first dt."""
        s = self.s
        s.update_ghost(do_swap=1)
        if s.num.enable_les > 0:
            s.compute_aux_les()
        else:
            s.compute_aux()
            jduc, jdiv = s.ns + 9, s.ns + 10
            s.be.call("bcextr_var", s.nx, s.ny, s.nz, s.ng, s.w_aux[jdiv - 1], s.be.stream)
            s.bcswap_var(jdiv)
            s.bcswap_var(jduc)
        s.compute_dt(self.c.cfl)

    # ------------------------------------------------------------------ restarts
    def field_info(self) -> dict:
        """The counters write_field_info stores (multideal.F90:2442-2468)."""
        return dict(icyc0=self.icyc, time0=self.s.time, itav=0, itslice_vtr=0, time_from_last_rst=0.0,
                    time_from_last_write=self.time_from_last_write, time_from_last_stat=0.0,
                    time_from_last_slice=0.0, time_from_last_slice_vtr=0.0, istore=self.istore)

    def write_restart(self, directory, io_type: int = 2):
        """The end of run: the field (rst.bin, or rst1_*.bin with io_type 1) and
        field_info0.dat (field_info1.dat with io_type 1)."""
        from . import restart
        self.s.save_field(directory, io_type)
        restart.write_field_info(pathlib.Path(directory) / ("field_info0.dat" if io_type == 2 else "field_info1.dat"),
                                 self.field_info())

    def read_restart(self, directory, io_type: int = 2):
        """restart_type 1: the field (rst.bin, or rst0_*.bin with io_type 1) and the
        counters of field_info0.dat; the run continues from icyc0 and time0."""
        from . import restart
        self.s.load_field(directory, io_type)
        info = restart.read_field_info(pathlib.Path(directory) / "field_info0.dat")
        self.icyc = self.icyc0 = int(info["icyc0"])
        self.s.time = float(info["time0"])
        self.time_from_last_write = float(info["time_from_last_write"])
        self.istore = int(info["istore"])
        return info

    # ------------------------------------------------------------------ output
    def _aux_host(self):
        return self.s.be.to_numpy(self.s.w_aux)

    def manage_output(self):
        """The writers of manage_output (multideal_gpu.F90:2465-2564) used by the
        validation cases."""
        s, c = self.s, self.c
        ns, ng = s.ns, s.ng
        self.time_from_last_write += s.dt
        if self.time_from_last_write >= c.dtsave:
            wa = self._aux_host()
            if self.out is not None:
                with open(self.out / "output_1d.dat", "w") as fh:
                    x = s.axes[0].coord
                    if s.nx > 1:
                        for i in range(s.nx + 2 * ng):
                            fh.write(FMT_1D.format(x[i]) + "".join(FMT_1D.format(v) for v in wa[:, ng, ng, i]) + "\n")
            self.time_from_last_write = 0.0
            self.istore += 1
        if c.flow_init == 8:
            nstep = int(round(s.time / s.dt))
            if nstep >= 10 and nstep % 10 == 0:
                wa = self._aux_host()
                row = (s.time, wa[ns + 10, ng, ng, ng], wa[ns + 4, ng, ng, ng])
                self.batch_rows.append(row)
                if self.out is not None:
                    with open(self.out / "output_batch.dat", "a" if len(self.batch_rows) > 1 else "w") as fh:
                        fh.write("{:20.10E}{:20.10E}{:20.10E}\n".format(*row))
        elif c.flow_init == 9:
            for q, (tm, name) in enumerate(((170e-6, 170), (190e-6, 190), (230e-6, 230))):
                if s.time >= tm and self.tube_done[q] == 0:
                    self.tube_done[q] = 1
                    wa = self._aux_host()
                    x = s.axes[0].coord[ng:-ng]
                    cols = np.stack([x, wa[ns + 4, ng, ng, ng:-ng], wa[ns, ng, ng, ng:-ng], wa[2, ng, ng, ng:-ng]], 1)
                    setattr(self, f"tube_{name}", cols)
                    if self.out is not None:
                        np.savetxt(self.out / f"output_1d_{name}micro_s.dat", cols, fmt="%20.10E", delimiter="")
                    break

    def print_progress(self, step_time: float):
        e = self.extrema
        row = (self.icyc, step_time, self.s.dt, self.s.time, self.residual_rhou, e["rhomin"], e["rhomax"],
               e["tmin"], e["tmax"], e["pmin"], e["pmax"])
        self.progress.append(row)
        if self.out is not None:
            mode = "w" if self.icyc == 1 else "a"
            with open(self.out / "progress.out", mode) as fh:
                fh.write("{:10d}{:18.10f}".format(row[0], row[1]) + "".join("{:20.10E}".format(v) for v in row[2:]) + "\n")

    # ------------------------------------------------------------------ input
    @classmethod
    def from_ini(cls, path, backend, mechanism: Optional[str] = None, outdir: Optional[str] = None,
                 nx: Optional[int] = None) -> "Run":
        """A run of the reference's ini (its directory holds input_cantera.yaml unless
        `mechanism` names the YAML); nx overrides nxmax."""
        ini = read_ini(path)
        ref, grid, bc = ini["ref_state"], ini["grid"], ini["bc"]
        ctl, num, flow, fluid = ini["controls"], ini["numerics"], ini["flow"], ini["fluid"]
        out = ini.get("output", {})
        les = ini.get("lespar", {})
        mech = Mechanism(mechanism or str(pathlib.Path(path).parent / "input_cantera.yaml"))
        y = np.array([float(ref.get(f"init_y_{k + 1:02d}", 0.0)) for k in range(len(mech.species))])
        rho0 = float(ref["rho0"]) if "rho0" in ref else None
        p0 = float(ref["p0"])
        t0 = float(ref["t0"]) if "t0" in ref else None
        tb = mech.tables(p0=p0, t_min_tab=_num(fluid, "t_min_tab", 200.0), t_max_tab=_num(fluid, "t_max_tab", 20000.0),
                         dt_tab=_num(fluid, "dt_tab", 1.0))
        if rho0 is None:
            rho0 = p0 / (tb.rmix(y) * t0)
        fp = cases.flow_params(tb, y, rho0, p0, float(ref.get("u0", 0.0)))
        numerics = solvermod.Numerics(
            ep_order=_num(num, "ep_order", 6, int), nkeep=_num(num, "nkeep", 0, int),
            visc_order=_num(num, "visc_order", 6, int),
            conservative_viscous=_num(num, "conservative_viscous", 1, int),
            weno_scheme=_num(num, "weno_scheme", 3, int), weno_version=_num(num, "weno_version", 1, int),
            flux_splitting=_num(num, "flux_splitting", 0, int),
            sensor_threshold=_num(num, "sensor_threshold", 0.1), sensor_type=_num(num, "sensor_type", 0, int),
            enable_chemistry=_num(fluid, "enable_chemistry", 0, int),
            correct_bound_ord=_num(bc, "correct_bound_ord", 0, int), ros_eps=_num(num, "ros_eps", 1e-8),
            ros_maxsteps=_num(num, "ros_maxsteps", 10000, int), ros_maxtry=_num(num, "ros_maxtry", 40, int),
            rosenbrock_version=_num(num, "rosenbrock_version", 1, int),
            operator_splitting=_num(num, "operator_splitting", 2, int))
        if _num(les, "enable_les", 0, int) > 0:
            # multideal.F90:615-646: the constants the model needs are required keys
            model = _num(les, "les_model", 0, int)
            keys = ["les_pr", "les_sc", "les_c_wale"] + (["les_c_yoshi", "les_c_eps", "les_c_mix"] if model == 2 else [])
            missing = [k for k in keys if k not in les]
            if missing:
                raise KeyError(f"[lespar] needs {', '.join(missing)}")
            numerics = dataclasses.replace(numerics, enable_les=1, les_model=model,
                                           **{k: float(les[k]) for k in keys})
        shape = [nx or _num(grid, "nxmax", 1, int), _num(grid, "nymax", 1, int), _num(grid, "nzmax", 1, int)]
        ng = _num(grid, "ng", 3, int)
        lengths = [_num(grid, "domain_size_x", 1.0), _num(grid, "domain_size_y", 1.0), _num(grid, "domain_size_z", 1.0)]
        tags = tuple(_num(bc, k, 0, int) for k in ("xmin", "xmax", "ymin", "ymax", "zmin", "zmax"))
        finit = _num(flow, "flow_init", -1, int)
        s = solvermod.Solver(backend, tb, *shape, ng, lengths, bctags=tags, numerics=numerics, rho0=fp["rho0"],
                             u0=fp["u0"], p0=p0)
        if finit == 9:
            w, tg, _ = cases.shock_tube(tb, shape[0], y, fp["rho0"], p0, ng=ng, length=lengths[0], ny=shape[1],
                                        nz=shape[2])
        elif finit == 8:
            w, tg = cases.reactor_state(tb, shape, ng, fp["t0"], p0, y)
        else:
            raise NotImplementedError(f"flow_init {finit}: only 8 (0-D reactor) and 9 (reactive tube) are set up here")
        s.set_state(w, tg)
        ctrl = Controls(cfl=_num(ctl, "cfl", 0.5), iter_dt_recompute=_num(ctl, "iter_dt_recompute", 1, int),
                        num_iter=_num(ctl, "num_iter", 1, int), print_control=_num(out, "print_control", 1, int),
                        dtsave=_num(out, "dtsave", 1e6), flow_init=finit)
        r = cls(s, ctrl, outdir)
        r.tables = tb
        return r

