"""Device-resident field and the reference's RK3 orchestration over the C-ABI.

``Solver`` mirrors the parts of ``equation_multideal_gpu_object`` on the hot path:
``rk_sync`` (multideal_gpu.F90:536-621), ``compute_aux`` (:361-443), ``euler_x/y/z``
(:948-1090), ``visflx`` (:302-359), ``compute_chemistry`` (:840-875), ``update_ghost``
(:1165-1238), ``compute_dt`` (:1620-1657) and the Strang-split implicit chemistry of
``advance_solution`` (:1853-1931).  Arrays keep the Fortran layouts of the reference
(stored as C-contiguous arrays with the axes reversed: ``w[m, k, j, i]``).

The numerics are executed by a *backend*: ``HipBackend`` (torch device memory + the
``libstreams_hip.so`` entry points) is the product path.  Any object with the same
``call/zeros/asarray/to_numpy/sync`` methods can drive the same orchestration (the tests
use one that runs the CPU restatement, to compare both sides step by step).
"""
from __future__ import annotations

import dataclasses
import pathlib
from typing import Optional, Sequence

import numpy as np

from . import capi
from . import grid as gridmod
from .mechanism import R_UNIV, Tables

TOL_ITER_NR = 1.0e-12   # src/parameters.F90:49
WRAY_RHO = (0.0, -17.0 / 60.0, -5.0 / 12.0)   # multideal.F90:5768-5773
WRAY_GAM = (8.0 / 15.0, 5.0 / 12.0, 3.0 / 4.0)


class HipBackend:
    """torch tensors on the GPU + libstreams_hip.so."""

    name = "hip"

    def __init__(self, device: str = "cuda:0"):
        import torch
        if not torch.cuda.is_available():
            raise capi.StreamsError("no GPU visible: the HIP backend has no CPU fallback")
        self.torch = torch
        self.device = torch.device(device)
        self.lib = capi.load()

    @property
    def stream(self):
        return self.torch.cuda.current_stream(self.device).cuda_stream

    def zeros(self, shape, dtype="f8"):
        t = self.torch
        dt = {"f8": t.float64, "i4": t.int32}[dtype]
        return t.zeros(shape, dtype=dt, device=self.device)

    def asarray(self, a: np.ndarray, dtype="f8"):
        t = self.torch
        a = np.ascontiguousarray(a, dtype={"f8": np.float64, "i4": np.int32}[dtype])
        return t.from_numpy(a).to(self.device)

    def to_numpy(self, x) -> np.ndarray:
        return x.detach().cpu().numpy()

    def copy_(self, dst, src):
        dst.copy_(src)

    def fill_(self, dst, value: float):
        dst.fill_(value)

    def sync(self):
        self.torch.cuda.synchronize(self.device)

    def call(self, name, *args):
        return self.lib.call(name, *args)

    def has(self, name) -> bool:
        return self.lib.has(name)


@dataclasses.dataclass
class Numerics:
    ep_order: int = 6
    nkeep: int = 0
    visc_order: int = 6
    conservative_viscous: int = 1
    weno_scheme: int = 3
    weno_version: int = 1
    flux_splitting: int = 0
    sensor_threshold: float = 2.0
    sensor_type: int = 1
    enable_chemistry: int = 1
    correct_bound_ord: int = 0
    ros_eps: float = 1e-8
    ros_maxsteps: int = 10000
    ros_maxtry: int = 40
    rosenbrock_version: int = 1      # 1: ROS4 (rosenbrock), 2: Rosenbrock-Krylov
    operator_splitting: int = 2      # with enable_chemistry == 2: 1 simpler (Krylov only), 2 Strang
    # with enable_chemistry == 2 and Strang splitting: -1 integrates each cell with the
    # reference's field kernel (rosenbrock_version); 0, 1, 2 integrate the interior cells as
    # one batch of constant-volume reactors (streams_reactor_batch form `mixed`: 0 FP64 in
    # the reference's arithmetic, 1 mixed precision, 2 FP64 species block)
    chem_batch_form: int = -1
    # [lespar] (multideal.F90:615-646): enable_les with les_model 1 (WALE) or 2 (WALE +
    # PaSR closure of the implicit chemistry); the constants have no defaults there
    enable_les: int = 0
    les_model: int = 1
    les_pr: float = 0.0
    les_sc: float = 0.0
    les_c_wale: float = 0.0
    les_c_yoshi: float = 0.0
    les_c_eps: float = 0.0
    les_c_mix: float = 0.0

    @property
    def enable_pasr(self) -> int:
        return int(self.enable_les > 0 and self.les_model == 2)


class Solver:
    """One rank's subdomain: padded field, auxiliaries, RHS buffers and tables on a backend."""

    def __init__(self, backend, tables: Tables, nx: int, ny: int, nz: int, ng: int,
                 lengths: Sequence[float], bctags: Sequence[int] = (0, 0, 0, 0, 0, 0),
                 numerics: Optional[Numerics] = None, rho0: float = 1.0, u0: float = 1.0,
                 p0: Optional[float] = None, eps_sensor: Optional[float] = None, mappings=None,
                 bctags_nr: Sequence[int] = (0, 0, 0, 0, 0, 0), winf: Optional[np.ndarray] = None):
        """mappings: optional per-axis callables x(xi) giving stretched node coordinates
        (metrics from the coordinates by the reference's central differences).
        bctags_nr: the characteristic treatment of each face (xmin_nr .. zmax_nr: 0 none,
        1 non-reflecting, 2 relaxation to the ghost cell, 3 to winf, 6 reflective wall);
        winf: the freestream conserved state (nv) that nr_type 3 relaxes to."""
        self.be = backend
        self.tb = tables
        self.ns = tables.n_species
        self.nv = self.ns + 4
        self.num = numerics or Numerics()
        num = self.num
        if num.enable_les > 0:
            if num.les_model not in (1, 2):
                raise ValueError("LES model not implemented")
            if num.les_model == 2 and num.enable_chemistry != 2:
                raise ValueError("PaSR cannot be used without chemistry enabled or for explicit integration")
        # nv_aux of multideal.F90:651-654: the eddy viscosity slot (J_LES1), the chemistry
        # diagnostics and the PaSR factor come after the 2 N_S + 13 base variables
        self.nv_aux = (2 * self.ns + 13 + (1 if num.enable_les > 0 else 0)
                       + (2 * (self.ns + 1) if num.enable_chemistry > 0 else 0) + num.enable_pasr)
        self.nx, self.ny, self.nz, self.ng = nx, ny, nz, ng
        self.bctags = list(bctags)
        self.periodic = np.array([int(bctags[0] == 0 and bctags[1] == 0), int(bctags[2] == 0 and bctags[3] == 0),
                                  int(bctags[4] == 0 and bctags[5] == 0)], dtype=np.int32)
        self.rho0, self.u0 = rho0, u0
        self.p0 = tables.p0 if p0 is None else p0
        self.eps_sensor = u0 * u0 if eps_sensor is None else eps_sensor
        maps = list(mappings) if mappings is not None else [None, None, None]
        self.axes = [gridmod.uniform_axis(n, L, ng, bool(self.periodic[d])) if maps[d] is None
                     else gridmod.mapped_axis(n, L, ng, bool(self.periodic[d]), maps[d])
                     for d, (n, L) in enumerate(zip((nx, ny, nz), lengths))]
        be = backend
        px, py, pz = nx + 2 * ng, ny + 2 * ng, nz + 2 * ng
        self.w = be.zeros((self.nv, pz, py, px))
        self.w_aux = be.zeros((self.nv_aux, pz, py, px))
        self.fl = be.zeros((self.nv, nz, ny, nx))
        self.fln = be.zeros((self.nv, nz, ny, nx))
        self.fhat = be.zeros((self.nv, pz, py, px))
        # fused substeps: fold each substep's init_flux into the previous substep's tail
        # within rk_sync (bit-identical; saves one read and one write of fl and fln)
        self.fold_substeps = True
        self._fln_set = False
        # tables (Fortran layouts: store transposed C arrays)
        f = lambda a: be.asarray(np.asarray(a, dtype=np.float64).T)  # noqa: E731
        self.t_rgas = be.asarray(tables.rgas)
        self.t_mw = be.asarray(tables.mw)
        self.t_mwinv = be.asarray(tables.mwinv)
        self.t_cp = f(tables.cp_coeff)
        self.t_cv = f(tables.cv_coeff)
        self.t_trange = f(tables.trange)
        self.t_visc = f(tables.visc_species)
        self.t_lambda = f(tables.lambda_species)
        self.t_diff = f(tables.diffbin_species)
        self.nr = tables.nreactions if self.num.enable_chemistry > 0 else 0
        if self.nr > 0:
            self.t_arr_a, self.t_arr_b, self.t_arr_ea = f(tables.arr_a), f(tables.arr_b), f(tables.arr_ea)
            self.t_fall, self.t_tb = f(tables.falloff_coeffs), f(tables.tb_eff)
            self.t_rco, self.t_pco, self.t_kc = f(tables.r_coeffs), f(tables.p_coeffs), f(tables.kc_tab)
            self.t_rty = be.asarray(tables.reac_ty, "i4")
            self.t_rev = be.asarray(tables.isRev, "i4")
        self.cd1 = f(gridmod.coeff_deriv1_table())
        self.cd2 = f(gridmod.coeff_deriv2_table())
        self.t_x, self.t_y, self.t_z = (be.asarray(a.coord) for a in self.axes)
        self.dxi = [be.asarray(a.dxi) for a in self.axes]
        self.dxi2 = [be.asarray(a.dxi2) for a in self.axes]
        self.dxi_2 = [be.asarray(a.dxi_2) for a in self.axes]
        epo = gridmod.ep_ord_change(nx, ny, nz, self.bctags, self.num.ep_order, self.num.weno_scheme,
                                    self.num.correct_bound_ord)
        self.t_epo = be.asarray(epo.T, "i4")
        self.t_periodic = self.periodic.copy()
        # Bilger mixture fraction off (N_EoI = 1)
        self.t_eoi = be.asarray(np.ones(1))
        self.t_nainsp = be.asarray(np.zeros(self.ns, dtype=np.int32), "i4")
        self.t_beta0 = be.asarray(np.zeros(2))
        # bc_preproc (multideal.F90:5671-5718): a face with a characteristic treatment is
        # left out of the convective update
        self.bctags_nr = list(bctags_nr)
        off = [1 if t != 0 else 0 for t in self.bctags_nr]
        self.eul = (1 + off[0], nx - off[1], 1 + off[2], ny - off[3], 1 + off[4], nz - off[5])
        self.t_winf = be.asarray(np.asarray(winf, dtype=np.float64)) if winf is not None else None
        self.force_zero_flux = [0] * 6
        self.dt = 0.0
        self.time = 0.0

    # ------------------------------------------------------------------ data
    def set_state(self, w_padded: np.ndarray, t_guess: np.ndarray):
        """w_padded[m,k,j,i] conserved variables incl. ghosts; t_guess the Newton start."""
        be = self.be
        be.copy_(self.w, be.asarray(w_padded))
        be.copy_(self.w_aux[self.ns + 4], be.asarray(t_guess))

    # ------------------------------------------------------------- restarts
    def save_field(self, directory, io_type: int = 2, ncoords=(0, 0, 0)):
        """write_field (io_type 2: rst.bin, field.F90:427-477) or write_field_serial
        (io_type 1: rst1_XXXX_XXXX_XXXX.bin, :412-425) of the conserved field."""
        from . import restart
        be, d = self.be, pathlib.Path(directory)
        d.mkdir(parents=True, exist_ok=True)
        if io_type == 2:
            dense = be.zeros((self.nv, self.nz, self.ny, self.nx))
            be.call("field_interior", self.nx, self.ny, self.nz, self.ng, self.nv, self.w, dense, 1, be.stream)
            be.sync()
            restart.write_rst_bin(d / "rst.bin", be.to_numpy(dense))
        else:
            ng = self.ng
            host = np.empty((self.nz + 2 * ng, self.ny + 2 * ng, self.nx + 2 * ng, self.nv))
            be.call("copy_gpu_cpu", self.nx, self.ny, self.nz, ng, self.nv, self.w, host, be.stream)
            restart.write_rst_serial(d / restart.serial_name("rst1", ncoords), host[ng:-ng, ng:-ng, ng:-ng, :])

    def load_field(self, directory, io_type: int = 2, ncoords=(0, 0, 0)):
        """read_field (rst.bin) or read_field_serial (rst0_XXXX_XXXX_XXXX.bin) into the
        interior of the conserved field (field.F90:349-410); the ghost layers keep their
        values until the boundary conditions and the halo exchange renew them."""
        from . import restart
        be, d = self.be, pathlib.Path(directory)
        if io_type == 2:
            a = restart.read_rst_bin(d / "rst.bin", self.nx, self.ny, self.nz, self.nv)
            be.call("field_interior", self.nx, self.ny, self.nz, self.ng, self.nv, self.w, be.asarray(a), 0,
                    be.stream)
            be.sync()
        else:
            ng = self.ng
            a = restart.read_rst_serial(d / restart.serial_name("rst0", ncoords), self.nx, self.ny, self.nz, self.nv)
            # the host copy keeps its ghost values, as the reference's host field does across
            # read_field_serial (which fills 1:nx, 1:ny, 1:nz only) and copy_cpu_gpu
            host = np.empty((self.nz + 2 * ng, self.ny + 2 * ng, self.nx + 2 * ng, self.nv))
            be.call("copy_gpu_cpu", self.nx, self.ny, self.nz, ng, self.nv, self.w, host, be.stream)
            host[ng:-ng, ng:-ng, ng:-ng, :] = a
            be.call("copy_cpu_gpu", self.nx, self.ny, self.nz, ng, self.nv, host, self.w, be.stream)

    # ------------------------------------------------------------- kernels
    def eval_aux(self, box=None):
        nx, ny, nz, ng = self.nx, self.ny, self.nz, self.ng
        if box is None:
            box = (1 - ng, nx + ng, 1 - ng, ny + ng, 1 - ng, nz + ng)
        tb = self.tb
        self.be.call("eval_aux", self.ns, nx, ny, nz, ng, self.nv, self.nv_aux, *box, self.w, self.w_aux,
                     self.p0, tb.t_min_tab, tb.dt_tab, R_UNIV, self.t_rgas, self.t_cv, self.t_cp, tb.nsetcv,
                     self.t_trange, tb.indx_cp_l, tb.indx_cp_r, TOL_ITER_NR, self.be.stream, self.t_mw,
                     self.t_mwinv, self.t_visc, self.t_lambda, self.t_diff, tb.num_t_tab, 1, self.t_eoi,
                     self.t_nainsp, self.t_beta0, self.t_eoi)

    def eval_aux2(self, ducros: bool = True):
        """eval_aux2_cuf; ducros=False (the fused substeps with the sensor off) computes
        div/3 only and leaves the Ducros column as it was."""
        self.be.call("eval_aux2", self.ns, self.nx, self.ny, self.nz, self.ng, self.nv_aux, self.num.visc_order,
                     self.w_aux, self.cd1, self.dxi[0], self.dxi[1], self.dxi[2], self.eps_sensor,
                     self.num.sensor_type if ducros else -1, self.be.stream)

    def compute_aux_les(self):
        """compute_aux_les (multideal_gpu.F90:445-498): Y, rho and velocity, the WALE eddy
        viscosity with the sensor, their ghost layers, then the thermodynamic state and the
        transport with the eddy viscosity added."""
        n, be = self.num, self.be
        tb = self.tb
        nx, ny, nz, ng = self.nx, self.ny, self.nz, self.ng
        be.call("eval_velaux", self.ns, nx, ny, nz, ng, self.nv, self.nv_aux, self.w, self.w_aux, be.stream)
        be.call("les_wale_mut", self.ns, nx, ny, nz, ng, self.nv_aux, n.ep_order, self.w_aux, self.cd1, self.dxi[0],
                self.dxi[1], self.dxi[2], n.les_c_wale, self.eps_sensor, n.sensor_type, be.stream)
        jduc, jdiv, jles = self.ns + 9, self.ns + 10, 2 * self.ns + 14
        self.bcswap_var(jduc)
        be.call("bcextr_var", nx, ny, nz, ng, self.w_aux[jdiv - 1], be.stream)
        self.bcswap_var(jdiv)
        be.call("bcextr_var", nx, ny, nz, ng, self.w_aux[jles - 1], be.stream)
        self.bcswap_var(jles)
        be.call("eval_aux_les", self.ns, nx, ny, nz, ng, self.nv, self.nv_aux, self.w, self.w_aux, self.p0,
                tb.t_min_tab, tb.dt_tab, R_UNIV, self.t_rgas, self.t_cv, self.t_cp, tb.nsetcv, self.t_trange,
                tb.indx_cp_l, tb.indx_cp_r, TOL_ITER_NR, self.t_mw, self.t_mwinv, self.t_visc, self.t_lambda,
                self.t_diff, tb.num_t_tab, n.les_pr, n.les_sc, be.stream)

    def compute_chem_aux(self):
        """compute_chem_aux (multideal_gpu.F90:500-534): production rates and heat release
        rate into w_aux (the output-time diagnostics)."""
        n, tb = self.num, self.tb
        if getattr(self, "t_h298", None) is None:
            self.t_h298 = self.be.asarray(tb.h298)
        self.be.call("eval_chem_aux", self.ns, self.nx, self.ny, self.nz, self.nv_aux, self.ng, self.nr, self.w_aux,
                     self.t_mw, self.t_arr_a, self.t_arr_b, self.t_arr_ea, self.t_tb, self.t_fall, self.t_rty,
                     self.t_rev, self.t_rco, self.t_pco, self.t_kc, tb.num_t_tab, tb.t_min_tab, tb.dt_tab, R_UNIV,
                     self.dxi[0], self.dxi[1], self.dxi[2], self.t_h298, n.enable_pasr, n.les_c_yoshi, n.les_c_mix,
                     n.les_c_eps, self.be.stream)

    def compute_aux(self, central: int = 1, ghost: int = 1, ducros: bool = True):
        """compute_aux (multideal_gpu.F90:361-443); with LES every call site of the
        reference calls compute_aux_les instead, whatever the box.  ducros=False: see
        eval_aux2."""
        if self.num.enable_les > 0:
            self.compute_aux_les()
            return
        nx, ny, nz, ng = self.nx, self.ny, self.nz, self.ng
        if central == 1 and ghost == 1:
            self.eval_aux()
        elif central == 1 and ghost == 0:
            self.eval_aux((1, nx, 1, ny, 1, nz))
        elif central == 0 and ghost == 1:
            for box in ((1 - ng, 0, 1 - ng, ny + ng, 1 - ng, nz + ng), (nx + 1, nx + ng, 1 - ng, ny + ng, 1 - ng, nz + ng),
                        (1 - ng, nx + ng, 1 - ng, 0, 1 - ng, nz + ng), (1 - ng, nx + ng, ny + 1, ny + ng, 1 - ng, nz + ng),
                        (1 - ng, nx + ng, 1 - ng, ny + ng, 1 - ng, 0), (1 - ng, nx + ng, 1 - ng, ny + ng, nz + 1, nz + ng)):
                self.eval_aux(box)
        if ghost == 1:
            self.eval_aux2(ducros)

    def bcswap(self):
        self.be.call("bcswap", self.nx, self.ny, self.nz, self.ng, self.nv, self.w, self.t_periodic, self.be.stream)

    def bcswap_var(self, m: int):
        """Exchange one auxiliary variable (1-based index m)."""
        self.be.call("bcswap", self.nx, self.ny, self.nz, self.ng, 1, self.w_aux[m - 1], self.t_periodic,
                     self.be.stream)

    def euler(self, d: int):
        tb, n = self.tb, self.num
        lmax = n.ep_order // 2
        args_common = (self.ns, self.nv, self.nv_aux, self.nx, self.ny, self.nz, self.ng)
        if d == 0:
            imin, imax = self.eul[0], self.eul[1]
            self.be.call("euler_x", *args_common, imin - 1, imax, lmax, n.nkeep, self.t_rgas, self.w_aux, self.cd1,
                         self.dxi[0], self.fhat, self.force_zero_flux[0], self.force_zero_flux[1], n.weno_scheme,
                         n.weno_version, n.sensor_threshold, 2 * n.weno_scheme, self.t_cv, self.t_cp, tb.nsetcv,
                         self.t_trange, tb.indx_cp_l, tb.indx_cp_r, self.t_epo, TOL_ITER_NR, self.rho0, self.u0,
                         n.flux_splitting, imin, imax, self.fl, self.be.stream)
        else:
            name = "euler_y" if d == 1 else "euler_z"
            lo, hi = self.eul[2 * d], self.eul[2 * d + 1]
            self.be.call(name, *args_common, lo, hi, lmax, n.nkeep, self.t_rgas, self.w_aux, self.fl, self.cd1,
                         self.dxi[d], self.fhat, self.force_zero_flux[2 * d], self.force_zero_flux[2 * d + 1],
                         n.weno_scheme, n.weno_version, n.sensor_threshold, 2 * n.weno_scheme, self.t_cv,
                         self.t_cp, tb.nsetcv, self.t_trange, tb.indx_cp_l, tb.indx_cp_r, self.t_epo, TOL_ITER_NR,
                         self.rho0, self.u0, n.flux_splitting, self.be.stream)

    def visflx(self, mode: int):
        tb = self.tb
        if mode == 1:
            for d, (name, coord) in enumerate((("visflx_x", self.t_x), ("visflx_y", self.t_y), ("visflx_z", self.t_z))):
                if (self.nx, self.ny, self.nz)[d] > 1:
                    self.be.call(name, self.ns, self.nx, self.ny, self.nz, self.nv, self.nv_aux, self.ng, coord,
                                 self.w_aux, self.fl, self.fhat, tb.indx_cp_l, tb.indx_cp_r, self.t_cp, tb.nsetcv,
                                 self.t_trange, self.t_rgas, R_UNIV, self.be.stream)
        elif mode == 2:
            self.be.call("visflx_reduced", self.ns, self.nx, self.ny, self.nz, self.ng, self.nv_aux,
                         self.num.visc_order, self.w_aux, self.fl, self.cd1, self.dxi[0], self.dxi[1], self.dxi[2],
                         self.be.stream)
        else:
            # the Laplacian form (visflx_cuf): momentum and energy, no species diffusion
            self.be.call("visflx", self.ns, self.nx, self.ny, self.nz, self.ng, self.nv_aux, self.num.visc_order,
                         self.w_aux, self.fl, self.cd1, self.cd2, *self.dxi, *self.dxi2, *self.dxi_2,
                         self.be.stream)

    def compute_chemistry(self):
        tb = self.tb
        self.be.call("compute_chemistry", self.ns, self.nx, self.ny, self.nz, self.nv, self.nv_aux, self.ng, self.nr,
                     self.w, self.w_aux, self.fl, self.t_mw, self.t_arr_a, self.t_arr_b, self.t_arr_ea, self.t_tb,
                     self.t_fall, self.t_rty, self.t_rev, self.t_rco, self.t_pco, self.t_kc, tb.num_t_tab,
                     tb.t_min_tab, tb.dt_tab, R_UNIV, tb.indx_cp_l, tb.indx_cp_r, self.t_cv, tb.nsetcv,
                     self.t_trange, TOL_ITER_NR, self.be.stream)

    def rosenbrock(self, dt_chem: float):
        tb, n = self.tb, self.num
        self.be.call("rosenbrock", self.ns, self.nv, self.nv_aux, self.nx, self.ny, self.nz, self.ng, self.w,
                     self.w_aux, self.fl, self.t_mw, self.nr, self.t_arr_a, self.t_arr_b, self.t_arr_ea, self.t_tb,
                     self.t_fall, self.t_rty, self.t_rev, self.t_rco, self.t_pco, self.t_kc, tb.num_t_tab,
                     tb.t_min_tab, tb.dt_tab, R_UNIV, TOL_ITER_NR, self.t_cv, tb.nsetcv, self.t_trange, tb.indx_cp_l,
                     tb.indx_cp_r, dt_chem, self.time, self.time + dt_chem, n.ros_maxsteps, n.ros_maxtry, n.ros_eps,
                     0, n.enable_pasr, n.les_c_yoshi, n.les_c_mix, n.les_c_eps, self.dxi[0], self.dxi[1], self.dxi[2],
                     self.be.stream)

    def chemistry_batch(self, dt_chem: float):
        """The implicit chemistry of one Strang half-step with every interior cell as one
        reactor of a batch (streams_reactor_batch, form Numerics.chem_batch_form): the
        conserved variables and the temperature guess are copied out of the padded field
        (streams_field_interior), integrated over dt_chem from the first trial step dt_chem
        as the field kernel rosenbrock_kernel does (kernels_gpu.F90:8928-9193), and copied
        back with the last temperature of each reactor."""
        n, be = self.num, self.be
        if n.enable_pasr:
            raise ValueError("the reactor batch has no PaSR closure: use chem_batch_form = -1")
        if be.name != "hip":
            raise ValueError("chem_batch_form >= 0 runs the HIP reactor batch; this backend has none")
        nb = self.nx * self.ny * self.nz
        if getattr(self, "_chem_batch", None) is None:
            from .reactors import ReactorBatchGPU
            rb = ReactorBatchGPU(be, self.tb, tol=n.ros_eps, maxsteps=n.ros_maxsteps, maxtry=n.ros_maxtry)
            self._chem_batch = (rb, be.zeros((self.nv, nb)), be.zeros((nb,)), be.zeros((nb,)), be.zeros((nb,)))
        rb, wb, tgb, dtb, teb = self._chem_batch
        jt = self.ns + 4                                   # J_T (0-based)
        be.call("field_interior", self.nx, self.ny, self.nz, self.ng, self.nv, self.w, wb, 1, be.stream)
        be.call("field_interior", self.nx, self.ny, self.nz, self.ng, 1, self.w_aux[jt], tgb, 1, be.stream)
        be.fill_(dtb, dt_chem)
        be.fill_(teb, dt_chem)
        rb.integrate(wb, tgb, dtb, teb, mixed=n.chem_batch_form)
        be.call("field_interior", self.nx, self.ny, self.nz, self.ng, self.nv, self.w, wb, 0, be.stream)
        be.call("field_interior", self.nx, self.ny, self.nz, self.ng, 1, self.w_aux[jt], tgb, 0, be.stream)

    def rosenbrock_krylov(self, dt_chem: float, fl_sav=None):
        """rosenbrock_krylov (multideal_gpu.F90:769-838); fl_sav given: simpler splitting,
        the transport forcing subtracted from every chemistry right-hand side."""
        tb, n = self.tb, self.num
        self.be.call("rosenbrock_krylov", self.ns, self.nv, self.nv_aux, self.nx, self.ny, self.nz, self.ng,
                     self.w, self.w_aux, self.fl, self.t_mw, self.nr, self.t_arr_a, self.t_arr_b, self.t_arr_ea,
                     self.t_tb, self.t_fall, self.t_rty, self.t_rev, self.t_rco, self.t_pco, self.t_kc, tb.num_t_tab,
                     tb.t_min_tab, tb.dt_tab, R_UNIV, TOL_ITER_NR, self.t_cv, tb.nsetcv, self.t_trange, tb.indx_cp_l,
                     tb.indx_cp_r, dt_chem, self.time, self.time + dt_chem, int(fl_sav is not None), fl_sav,
                     n.ros_maxsteps, n.enable_pasr, n.les_c_yoshi, n.les_c_mix, n.les_c_eps, self.dxi[0], self.dxi[1],
                     self.dxi[2], self.be.stream)

    def update_ghost(self, do_swap: int = 0):
        """update_ghost (multideal_gpu.F90:1165-1238): the physical boundary conditions of
        the non-periodic faces, then with do_swap = 1 the halo exchange (the RK substeps
        call it with do_swap = 0)."""
        for ilat, tag in enumerate(self.bctags, start=1):
            if tag == 0:
                continue
            if tag == 2:
                self.be.call("bcextr", ilat, self.nx, self.ny, self.nz, self.ng, self.nv, self.w, self.be.stream)
            elif tag == 8:
                self.be.call("bcwall_ad", self.ns, ilat, self.nx, self.ny, self.nz, self.ng, self.w, self.be.stream)
            else:
                raise NotImplementedError(f"boundary tag {tag}")
        if do_swap == 1:
            self.bcswap()

    # -------------------------------------------------------------- RK step
    def rk_substep(self, istep: int, simpler: bool = False):
        """One substep of rk_sync (multideal_gpu.F90:562-618); simpler: rk_sync with
        simpler_splitting = 1, whose stages accumulate fl - fl_sav."""
        n = self.num
        if self._fln_set:
            raise RuntimeError("the previous fused substep already did this substep's init_flux: "
                               "continue with rk_substep_fused")
        rhodt = WRAY_RHO[istep] * self.dt
        gamdt = WRAY_GAM[istep] * self.dt
        be = self.be
        if simpler:
            be.call("init_flux_simpler", self.nx, self.ny, self.nz, self.nv, self.fl, self.fln, self.fl_sav, rhodt,
                    be.stream)
        else:
            be.call("init_flux", self.nx, self.ny, self.nz, self.nv, self.fl, self.fln, rhodt, be.stream)
        self._transport_rhs()
        if n.enable_chemistry == 1:
            self.compute_chemistry()
        if simpler:
            be.call("update_simpler_flux", self.nx, self.ny, self.nz, self.nv, self.fl, self.fln, self.fl_sav, gamdt,
                    be.stream)
        else:
            be.call("update_flux", self.nx, self.ny, self.nz, self.nv, self.fl, self.fln, gamdt, be.stream)
        be.call("update_field", self.nx, self.ny, self.nz, self.ng, self.nv, self.w, self.fln, None, be.stream)
        self.update_ghost()

    def sav_flx(self):
        """sav_flx (multideal_gpu.F90:623-692): the right-hand side at the current state,
        without the field update, kept in fl_sav as the chemistry's transport forcing."""
        be, n = self.be, self.num
        if getattr(self, "fl_sav", None) is None:
            self.fl_sav = be.zeros((self.nv, self.nz, self.ny, self.nx))
        # the reference's sav_flx passes an unset rhodt: fln is overwritten by the first
        # init_flux_simpler of the following rk_sync, so its value does not matter
        be.call("init_flux", self.nx, self.ny, self.nz, self.nv, self.fl, self.fln, 0.0, be.stream)
        self._transport_rhs()
        if n.enable_chemistry == 1:
            self.compute_chemistry()
        self.update_ghost()
        be.call("sav_flx", self.nx, self.ny, self.nz, self.nv, self.fl, self.fl_sav, be.stream)

    def _transport_rhs(self):
        """bcswap, compute_aux, sensor exchange, euler_x/y/z and visflx into fl."""
        n, be = self.num, self.be
        self.bcswap()
        if n.enable_les > 0:
            self.compute_aux_les()
        else:
            self.compute_aux()
            jduc, jdiv = self.ns + 9, self.ns + 10
            be.call("bcextr_var", self.nx, self.ny, self.nz, self.ng, self.w_aux[jdiv - 1], be.stream)
            self.bcswap_var(jduc)
            self.bcswap_var(jdiv)
        self.euler(0)
        if self.ny > 1:
            self.euler(1)
        if self.nz > 1:
            self.euler(2)
        if n.conservative_viscous == 1:
            self.visflx(1)
            self.visflx(2)
        else:
            self.visflx(0)
        self.bc_nr()

    def bc_nr(self):
        """bc_nr (multideal_gpu.F90:1116-1163): the characteristic treatment of the faces
        that have one."""
        tb = self.tb
        for ilat, nr_type in enumerate(self.bctags_nr, start=1):
            if nr_type == 0:
                continue
            d = (ilat - 1) // 2
            self.be.call("bc_nr_lat", d + 1, (ilat - 1) % 2 + 1, nr_type, self.ns, self.nx, self.ny, self.nz, self.ng,
                         self.nv, self.w_aux, self.w, self.fl, self.dxi[d], tb.indx_cp_l, tb.indx_cp_r, self.t_rgas,
                         self.t_cv, self.t_cp, tb.nsetcv, self.t_trange, self.t_winf, self.be.stream)

    def fused_supported(self) -> bool:
        """The fused substep covers full-range sweeps (no non-reflecting offsets), the
        conservative viscous form without LES; explicit chemistry is folded into the tail,
        implicit chemistry (enable_chemistry = 2) runs around the substeps."""
        return (self.eul == (1, self.nx, 1, self.ny, 1, self.nz) and self.num.conservative_viscous == 1
                and self.num.enable_les == 0
                and self.be.has("rhs_sweep") and self.be.has("rhs_finish"))

    def rk_substep_fused(self, istep: int, fold_next: bool = False):
        """One RK substep in five passes: halo + aux, sensor, fused x/y/z sweeps (with
        init_flux folded into the first), and the pointwise tail (reduced viscous terms,
        chemistry, update_flux, update_field), then the physical boundaries.  fold_next (not
        for the last substep): the tail also does substep istep + 1's init_flux
        (streams_rhs_finish_next), so fl is left unset and that substep's first sweep skips
        reading it."""
        n, tb, be = self.num, self.tb, self.be
        fold_next = fold_next and istep < 2 and be.has("rhs_finish_next")
        rhodt = WRAY_RHO[istep] * self.dt
        gamdt = WRAY_GAM[istep] * self.dt
        self.bcswap()
        # the Ducros sensor is read only by the sweeps' shock test, which a threshold of 1 or
        # more (the sensor's upper bound) switches off
        sensor = n.sensor_threshold < 1.0
        self.compute_aux(ducros=sensor)
        jduc, jdiv = self.ns + 9, self.ns + 10
        be.call("bcextr_var", self.nx, self.ny, self.nz, self.ng, self.w_aux[jdiv - 1], be.stream)
        if sensor:
            self.bcswap_var(jduc)
        self.bcswap_var(jdiv)
        lmax = n.ep_order // 2
        first = True
        for d, (nd, coord) in enumerate(zip((self.nx, self.ny, self.nz), (self.t_x, self.t_y, self.t_z))):
            if d > 0 and nd <= 1:
                continue
            be.call("rhs_sweep", d + 1, self.ns, self.nv_aux, self.nx, self.ny, self.nz, self.ng, lmax, n.nkeep,
                    self.t_rgas, self.w_aux, self.cd1, self.dxi[d], coord, self.fhat, self.force_zero_flux[2 * d],
                    self.force_zero_flux[2 * d + 1], n.weno_scheme, n.weno_version, n.sensor_threshold, self.t_cv,
                    self.t_cp, tb.nsetcv, self.t_trange, tb.indx_cp_l, tb.indx_cp_r, self.t_epo, TOL_ITER_NR,
                    self.rho0, self.u0, n.flux_splitting, int(nd > 1), R_UNIV,
                    (2 if self._fln_set else 1) if first else 0, rhodt, self.fl, self.fln, be.stream)
            first = False
        chem = int(n.enable_chemistry == 1 and self.nr > 0)
        if chem and getattr(self, "t_kinpack", None) is None:
            self.t_kinpack = be.asarray(capi.kinetics_pack(tb, be.lib))
        kin = (self.t_mw, self.t_kinpack, self.t_kc) if chem else (self.t_mw, None, None)
        args = (self.ns, self.nx, self.ny, self.nz, self.ng, self.nv_aux, n.visc_order, self.w, self.w_aux, self.fl,
                self.fln, self.cd1, self.dxi[0], self.dxi[1], self.dxi[2], 1, chem, self.nr, *kin, tb.num_t_tab,
                tb.t_min_tab, tb.dt_tab, R_UNIV, gamdt)
        if fold_next:
            be.call("rhs_finish_next", *args, WRAY_RHO[istep + 1] * self.dt, None, be.stream)
        else:
            be.call("rhs_finish", *args, None, be.stream)
        self._fln_set = fold_next
        self.update_ghost()

    def rk_sync(self, fused: Optional[bool] = None, simpler: bool = False):
        """rk_sync (multideal_gpu.F90:536-621).  The simpler-splitting stages run routine
        by routine (the fused passes fold update_flux, not its simpler variant)."""
        if fused is None:
            fused = self.fused_supported()
        for istep in range(3):
            if fused and not simpler:
                self.rk_substep_fused(istep, fold_next=self.fold_substeps)
            else:
                self.rk_substep(istep, simpler=simpler)

    def advance(self, fused: Optional[bool] = None):
        """advance_solution (multideal_gpu.F90:1853-1931) without energy deposition."""
        n = self.num
        if n.enable_chemistry == 2 and n.operator_splitting == 1:
            if n.rosenbrock_version != 2:
                raise ValueError("simpler splitting needs the Rosenbrock-Krylov integrator (rosenbrock_version 2)")
            if n.chem_batch_form >= 0:
                raise ValueError("the reactor batch takes no transport forcing: simpler splitting needs "
                                 "chem_batch_form = -1")
            dt = self.dt
            self.dt = 0.5 * dt
            self.sav_flx()
            self.compute_aux(central=1, ghost=0)
            self.rosenbrock_krylov(dt, fl_sav=self.fl_sav)
            self.rk_sync(fused=False, simpler=True)
            self.dt = dt
        elif n.enable_chemistry == 2:
            dt_chem = 0.5 * self.dt
            if n.chem_batch_form >= 0:
                chem = self.chemistry_batch
            else:
                chem = self.rosenbrock if n.rosenbrock_version == 1 else self.rosenbrock_krylov
            chem(dt_chem)
            self.rk_sync(fused=fused)
            self.compute_aux(central=1, ghost=0)
            chem(dt_chem)
        else:
            self.rk_sync(fused=fused)
        self.time += self.dt

    def compute_dt(self, cfl: float) -> float:
        """compute_dt (multideal_gpu.F90:1620-1657)."""
        if cfl < 0:
            self.dt = -cfl
            return self.dt
        tb = self.tb
        out = np.zeros(9)
        self.be.call("compute_dt", self.ns, self.nx, self.ny, self.nz, self.ng, self.nv, self.nv_aux, self.dxi[0],
                     self.dxi[1], self.dxi[2], self.dxi2[0], self.dxi2[1], self.dxi2[2], self.w, self.w_aux, out,
                     tb.indx_cp_l, tb.indx_cp_r, self.t_cp, tb.nsetcv, self.t_trange, None, self.be.stream)
        self.dt = cfl / float(out.max())
        return self.dt
