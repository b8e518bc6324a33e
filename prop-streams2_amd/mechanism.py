"""Mechanism → kinetic / thermodynamic / transport tables in the reference's table layouts.

The multideal solver consumes a fixed set of tables built once at start-up
(``src/multideal/multideal.F90:2930-3318`` for the Cantera path, ``:2714-2928`` for the
``fluid_prop.bin`` path).  This module builds the same tables from a Cantera-format YAML
mechanism without Cantera: it parses the subset of the YAML schema that the reference's
example mechanisms use (NASA7/NASA9 thermo, elementary / three-body / falloff
Lindemann / Troe / SRI reactions, efficiencies, duplicates, irreversible reactions,
cm/mol/cal|kcal|K units) and evaluates

* NASA polynomials scaled by the species gas constant, ``cp_coeff``/``cv_coeff``
  (``multideal.F90:3036-3058``): index ``indx_cp_l..indx_cp_r`` are the c_p polynomial
  terms, ``indx_cp_r+1`` the enthalpy constant, ``indx_cp_r+2`` the entropy constant;
  ``cv_coeff(0)`` = ``cp_coeff(0) - R_s``.
* Arrhenius / falloff parameters in SI-kmol units (``cantera_utils_cxx.cxx:157-262``),
  Troe without T2 stored as the sentinel -3.14 (``:224-229``), SRI d/e defaults 1/0.
* Equilibrium constants in concentration units, K_c = exp(-ΔG°/RT)·(p°/RT)^Δν with
  p° = 1 atm, clamped to [tiny, huge] (``multideal.F90:3186-3194``).
* Pure-species viscosity / conductivity and binary diffusivities (at the reference
  pressure p0) tabulated on T = t_min_tab + (l-1)·dt_tab (``multideal.F90:3139-3167``)
  from Chapman–Enskog kinetic theory.  The reference takes these from Cantera's
  mixture-averaged transport model; Cantera is not available here, so the collision
  integrals come from the Neufeld et al. (1972) correlations with the Brokaw (1969)
  polar correction, and the conductivity from the Warnatz/Eucken-type split into
  translational, rotational and vibrational parts with Parker's rotational relaxation.
  These transport tables are not pinned against Cantera; GPU and CPU paths consume the
  same arrays, so kernel parity does not depend on them.

All arrays are returned Fortran-ordered with the reference's shapes so they can be handed
to the kernels through the C-ABI unchanged.
"""
from __future__ import annotations

import dataclasses
import math
import re
import struct
from typing import Dict, List, Optional, Sequence

import numpy as np
import yaml

R_UNIV = 8314.46261815324      # J/(kmol K), src/parameters.F90:54
ONE_ATM = 101325.0
NAME_LEN = 30                  # character(len=30) species_names, multideal.F90:213
AVOGADRO = 6.02214076e26       # 1/kmol
BOLTZMANN = 1.380649e-23       # J/K
EPS0 = 8.8541878128e-12        # F/m
LIGHT_SPEED = 299792458.0
DEBYE = 1.0e-21 / LIGHT_SPEED  # C m

ATOMIC_WEIGHTS = {  # kg/kmol (IUPAC values as used by Cantera's element table)
    "H": 1.008, "C": 12.011, "N": 14.007, "O": 15.999, "Ar": 39.95, "AR": 39.95,
    "He": 4.002602, "HE": 4.002602, "E": 5.48579909e-4,
}

REAC_ELEMENTARY, REAC_THREEBODY, REAC_LINDEMANN, REAC_TROE, REAC_SRI = 0, 1, 2, 3, 4
TROE_NO_T2 = -3.14


@dataclasses.dataclass
class Reaction:
    equation: str
    reactants: Dict[str, float]
    products: Dict[str, float]
    reversible: bool
    kind: int
    rate: Sequence[float]            # (A, b, Ea) high-pressure / only rate, SI
    low_rate: Optional[Sequence[float]] = None
    falloff: Sequence[float] = (0.0, 0.0, 0.0, 0.0, 0.0)
    efficiencies: Optional[Dict[str, float]] = None
    default_efficiency: float = 1.0


@dataclasses.dataclass
class Species:
    name: str
    composition: Dict[str, float]
    mw: float
    trange: List[float]
    coeffs: List[List[float]]        # per range, NASA7 (7) or NASA9 (9) coefficients
    geometry: int = 0                # 0 atom, 1 linear, 2 nonlinear
    well_depth: float = 0.0          # eps/kB [K]
    diameter: float = 0.0            # [Angstrom]
    dipole: float = 0.0              # [Debye]
    polarizability: float = 0.0      # [Angstrom^3]
    rot_relax: float = 0.0


# --------------------------------------------------------------------------------------
# YAML parsing
# --------------------------------------------------------------------------------------

def _unit_factors(units: dict):
    length = units.get("length", "m")
    quantity = units.get("quantity", "kmol")
    ea = units.get("activation-energy", "J/kmol")
    len_m = {"m": 1.0, "cm": 1e-2, "mm": 1e-3}[length]
    qty_kmol = {"kmol": 1.0, "mol": 1e-3, "molec": 1.0 / AVOGADRO}[quantity]
    conc = len_m ** 3 / qty_kmol        # converts (length^3/quantity) to m^3/kmol
    ea_fac = {
        "J/kmol": 1.0, "J/mol": 1e3, "kJ/mol": 1e6, "cal/mol": 4184.0,
        "kcal/mol": 4.184e6, "K": R_UNIV, "eV": 1.602176634e-19 * AVOGADRO,
    }[ea]
    return conc, ea_fac


_TERM = re.compile(r"^\s*(?:(\d+(?:\.\d*)?)\s+)?(\S+)\s*$")


def _parse_side(side: str):
    """Split one side of a reaction equation into species stoichiometry and collider."""
    collider = None
    m = re.search(r"\(\+\s*([^)]+?)\s*\)", side)
    if m:
        collider = m.group(1).strip()
        side = side[: m.start()] + side[m.end():]
    stoich: Dict[str, float] = {}
    third = False
    for term in side.split(" + "):
        term = term.strip()
        if not term:
            continue
        mt = _TERM.match(term)
        if not mt:
            raise ValueError(f"cannot parse reaction term {term!r}")
        coef = float(mt.group(1)) if mt.group(1) else 1.0
        name = mt.group(2)
        if name == "M":
            third = True
            continue
        stoich[name] = stoich.get(name, 0.0) + coef
    return stoich, collider, third


def _parse_equation(eq: str):
    for arrow, rev in (("<=>", True), ("=>", False), ("=", True)):
        if arrow in eq:
            lhs, rhs = eq.split(arrow, 1)
            break
    else:
        raise ValueError(f"no arrow in equation {eq!r}")
    r, rc, rthird = _parse_side(" " + lhs.strip() + " ")
    p, pc, pthird = _parse_side(" " + rhs.strip() + " ")
    return r, p, rev, rc or pc, rthird or pthird


def _rate(d: dict, order: float, conc: float, ea_fac: float):
    a = float(d["A"]) * conc ** (order - 1.0)
    return (a, float(d.get("b", 0.0)), float(d.get("Ea", 0.0)) * ea_fac)


class Mechanism:
    """A parsed mechanism restricted to a species subset, with table builders."""

    def __init__(self, path: str, species: Optional[Sequence[str]] = None, phase: int = 0,
                 reactions: bool = True):
        """species: a subset of the phase's species (every kept reaction must use only
        these; third-body efficiencies of dropped species are ignored).  reactions=False
        keeps the thermo and transport of the subset only (an inert mixture)."""
        with open(path) as fh:
            doc = yaml.safe_load(fh)
        self.path = path
        conc, ea_fac = _unit_factors(doc.get("units", {}))
        ph = doc["phases"][phase]
        all_names = list(ph["species"])
        names = list(species) if species is not None else all_names
        missing = [s for s in names if s not in all_names]
        if missing:
            raise ValueError(f"species {missing} not in mechanism phase")
        sdefs = {s["name"]: s for s in doc["species"]}
        self.species: List[Species] = [self._species(sdefs[n]) for n in names]
        self.names = names
        idx = {n: i for i, n in enumerate(names)}
        self.reactions: List[Reaction] = []
        dropped = set(all_names) - set(names)
        for rd in (doc.get("reactions", []) if reactions else []):
            rx = self._reaction(rd, conc, ea_fac)
            used = set(rx.reactants) | set(rx.products)
            if used & dropped:
                raise ValueError(f"reaction {rx.equation!r} uses a dropped species")
            if rx.efficiencies is not None:
                rx.efficiencies = {k: v for k, v in rx.efficiencies.items() if k in idx}
            self.reactions.append(rx)
        self.index = idx

    # ---------------------------------------------------------------- species
    @staticmethod
    def _species(sd: dict) -> Species:
        comp = {k: float(v) for k, v in sd["composition"].items()}
        mw = sum(ATOMIC_WEIGHTS[e] * n for e, n in comp.items())
        th = sd["thermo"]
        model = th["model"]
        if model == "NASA7":
            tr = [float(t) for t in th["temperature-ranges"]]
            data = [[float(c) for c in row] for row in th["data"]]
            if len(tr) == 2:
                # a single-range NASA7 entry is held as a two-range polynomial with
                # T_mid = T_max and the data repeated, as Cantera's NasaPoly2 does
                tr = [tr[0], tr[1], tr[1]]
                data = [data[0], data[0]]
            elif len(tr) != len(data) + 1:
                raise ValueError("NASA7 ranges / data mismatch")
        elif model == "NASA9":
            tr = [float(t) for t in th["temperature-ranges"]]
            data = [[float(c) for c in row] for row in th["data"]]
        else:
            raise ValueError(f"thermo model {model} not supported")
        tp = sd.get("transport", {})
        geo = {"atom": 0, "linear": 1, "nonlinear": 2}[tp.get("geometry", "atom")]
        return Species(
            name=sd["name"], composition=comp, mw=mw, trange=tr, coeffs=data, geometry=geo,
            well_depth=float(tp.get("well-depth", 0.0)), diameter=float(tp.get("diameter", 0.0)),
            dipole=float(tp.get("dipole", 0.0)), polarizability=float(tp.get("polarizability", 0.0)),
            rot_relax=float(tp.get("rotational-relaxation", 0.0)),
        )

    # --------------------------------------------------------------- reactions
    @staticmethod
    def _reaction(rd: dict, conc: float, ea_fac: float) -> Reaction:
        eq = rd["equation"]
        reac, prod, rev, collider, third = _parse_equation(eq)
        rtype = rd.get("type", "elementary")
        order = sum(reac.values())
        effs = None
        deff = 1.0
        if rtype == "elementary":
            if third:
                rtype = "three-body"
        if rtype in ("three-body", "falloff"):
            effs = {k: float(v) for k, v in rd.get("efficiencies", {}).items()}
            deff = float(rd.get("default-efficiency", 1.0))
            if collider is not None and collider != "M":
                effs = {collider: 1.0}
                deff = 0.0
        if rtype == "elementary":
            return Reaction(eq, reac, prod, rev, REAC_ELEMENTARY,
                            _rate(rd["rate-constant"], order, conc, ea_fac))
        if rtype == "three-body":
            return Reaction(eq, reac, prod, rev, REAC_THREEBODY,
                            _rate(rd["rate-constant"], order + 1.0, conc, ea_fac),
                            efficiencies=effs, default_efficiency=deff)
        if rtype == "falloff":
            hi = _rate(rd["high-P-rate-constant"], order, conc, ea_fac)
            lo = _rate(rd["low-P-rate-constant"], order + 1.0, conc, ea_fac)
            if "Troe" in rd:
                t = rd["Troe"]
                fc = [float(t["A"]), float(t["T3"]), float(t["T1"]),
                      float(t["T2"]) if "T2" in t else TROE_NO_T2, 0.0]
                kind = REAC_TROE
            elif "SRI" in rd:
                s = rd["SRI"]
                fc = [float(s["A"]), float(s["B"]), float(s["C"]),
                      float(s.get("D", 1.0)), float(s.get("E", 0.0))]
                kind = REAC_SRI
            else:
                fc = [0.0] * 5
                kind = REAC_LINDEMANN
            return Reaction(eq, reac, prod, rev, kind, hi, low_rate=lo, falloff=fc,
                            efficiencies=effs, default_efficiency=deff)
        raise ValueError(f"reaction type {rtype!r} not supported ({eq})")

    # ------------------------------------------------------------------ sizes
    @property
    def n_species(self) -> int:
        return len(self.species)

    @property
    def n_reactions(self) -> int:
        return len(self.reactions)

    # -------------------------------------------------------------- thermo
    def thermo_tables(self):
        """cp/cv coefficient and range tables (multideal.F90:3007-3058)."""
        ncoef = {len(s.coeffs[0]) for s in self.species}
        if len(ncoef) != 1:
            raise ValueError("mixed NASA7/NASA9 not supported")
        ncoef = ncoef.pop()
        nsets = {len(s.coeffs) for s in self.species}
        if len(nsets) != 1:
            raise ValueError("different number of temperature ranges not supported")
        nsetcv = nsets.pop()
        if ncoef == 7:
            lo, hi = 0, 4
        elif ncoef == 9:
            lo, hi = -2, 4
        else:
            raise ValueError("cp model not supported")
        ns = self.n_species
        ncol = hi + 2 - lo + 1
        trange = np.zeros((ns, nsetcv + 1), order="F")
        cp = np.zeros((ncol, ns, nsetcv), order="F")
        rgas = np.array([R_UNIV / s.mw for s in self.species])
        for k, s in enumerate(self.species):
            if len(s.trange) != nsetcv + 1:
                raise ValueError(f"species {s.name}: range table length")
            trange[k, :] = s.trange
            for m in range(nsetcv):
                cp[:, k, m] = np.asarray(s.coeffs[m], dtype=np.float64) * rgas[k]
        cv = cp.copy(order="F")
        cv[0 - lo, :, :] = cp[0 - lo, :, :] - rgas[:, None]
        return dict(indx_cp_l=lo, indx_cp_r=hi, nsetcv=nsetcv, trange=trange,
                    cp_coeff=cp, cv_coeff=cv, rgas=rgas,
                    mw=np.array([s.mw for s in self.species]))

    def _range_index(self, k: int, T: np.ndarray) -> np.ndarray:
        tr = np.asarray(self.species[k].trange)
        nset = len(tr) - 1
        # Numerical-Recipes locate on the range boundaries, clamped to [1, nset]
        j = np.searchsorted(tr, T, side="right")
        return np.clip(j, 1, nset) - 1

    def h_RT_s_R(self, k: int, T: np.ndarray):
        """Dimensionless standard enthalpy h/RT and entropy s/R of species k."""
        T = np.asarray(T, dtype=np.float64)
        sp = self.species[k]
        c = np.asarray(sp.coeffs, dtype=np.float64)[self._range_index(k, T)]
        if c.shape[-1] == 7:
            a0, a1, a2, a3, a4, a5, a6 = (c[..., i] for i in range(7))
            h = a0 + a1 * T / 2 + a2 * T ** 2 / 3 + a3 * T ** 3 / 4 + a4 * T ** 4 / 5 + a5 / T
            s = a0 * np.log(T) + a1 * T + a2 * T ** 2 / 2 + a3 * T ** 3 / 3 + a4 * T ** 4 / 4 + a6
        else:
            a = [c[..., i] for i in range(9)]
            h = (-a[0] / T ** 2 + a[1] * np.log(T) / T + a[2] + a[3] * T / 2 + a[4] * T ** 2 / 3
                 + a[5] * T ** 3 / 4 + a[6] * T ** 4 / 5 + a[7] / T)
            s = (-a[0] / T ** 2 / 2 - a[1] / T + a[2] * np.log(T) + a[3] * T + a[4] * T ** 2 / 2
                 + a[5] * T ** 3 / 3 + a[6] * T ** 4 / 4 + a[8])
        return h, s

    def cp_R(self, k: int, T: np.ndarray):
        T = np.asarray(T, dtype=np.float64)
        c = np.asarray(self.species[k].coeffs, dtype=np.float64)[self._range_index(k, T)]
        if c.shape[-1] == 7:
            return c[..., 0] + c[..., 1] * T + c[..., 2] * T ** 2 + c[..., 3] * T ** 3 + c[..., 4] * T ** 4
        return (c[..., 0] / T ** 2 + c[..., 1] / T + c[..., 2] + c[..., 3] * T + c[..., 4] * T ** 2
                + c[..., 5] * T ** 3 + c[..., 6] * T ** 4)

    # --------------------------------------------------------------- kinetics
    def kinetics_tables(self, t_tab: np.ndarray):
        ns, nr = self.n_species, self.n_reactions
        reac_ty = np.zeros(nr, dtype=np.int32)
        is_rev = np.zeros(nr, dtype=np.int32)
        arr_a = np.zeros((nr, 2), order="F")
        arr_b = np.zeros((nr, 2), order="F")
        arr_ea = np.zeros((nr, 2), order="F")
        falloff = np.zeros((nr, 5), order="F")
        tb_eff = np.ones((nr, ns), order="F")
        r_co = np.zeros((nr, ns), order="F")
        p_co = np.zeros((nr, ns), order="F")
        for i, rx in enumerate(self.reactions):
            reac_ty[i] = rx.kind
            is_rev[i] = 1 if rx.reversible else 0
            if rx.kind in (REAC_ELEMENTARY, REAC_THREEBODY):
                arr_a[i, 0], arr_b[i, 0], arr_ea[i, 0] = rx.rate
            else:
                arr_a[i, 0], arr_b[i, 0], arr_ea[i, 0] = rx.low_rate
                arr_a[i, 1], arr_b[i, 1], arr_ea[i, 1] = rx.rate
                falloff[i, :] = rx.falloff
            if rx.efficiencies is not None:
                for k, name in enumerate(self.names):
                    tb_eff[i, k] = rx.efficiencies.get(name, rx.default_efficiency)
            for name, nu in rx.reactants.items():
                r_co[i, self.index[name]] += nu
            for name, nu in rx.products.items():
                p_co[i, self.index[name]] += nu
        # equilibrium constants in concentration units (kmol/m^3)^dnu
        t_tab = np.asarray(t_tab, dtype=np.float64)
        g_rt = np.empty((ns, t_tab.size))
        for k in range(ns):
            h, s = self.h_RT_s_R(k, t_tab)
            g_rt[k] = h - s
        dnu = (p_co - r_co)
        dg = dnu @ g_rt                                   # (nr, nT)
        log_kc = -dg + dnu.sum(axis=1)[:, None] * np.log(ONE_ATM / (R_UNIV * t_tab))[None, :]
        with np.errstate(over="ignore", under="ignore"):
            kc = np.exp(log_kc)
        tiny, huge = np.finfo(np.float64).tiny, np.finfo(np.float64).max
        kc = np.clip(kc, tiny, huge)
        kc_tab = np.asfortranarray(kc.T)                   # (num_t_tab+1, nr)
        return dict(reac_ty=reac_ty, isRev=is_rev, arr_a=arr_a, arr_b=arr_b, arr_ea=arr_ea,
                    falloff_coeffs=falloff, tb_eff=tb_eff, r_coeffs=r_co, p_coeffs=p_co,
                    kc_tab=kc_tab)

    # --------------------------------------------------------------- transport
    def transport_tables(self, t_tab: np.ndarray, p0: float):
        """Pure-species μ, λ and binary D_jk(T) at p0 on the table grid."""
        T = np.asarray(t_tab, dtype=np.float64)
        ns = self.n_species
        sig = np.array([s.diameter for s in self.species]) * 1e-10
        eps = np.array([s.well_depth for s in self.species])            # eps/kB
        dip = np.array([s.dipole for s in self.species]) * DEBYE
        alpha = np.array([s.polarizability for s in self.species]) * 1e-30
        mass = np.array([s.mw for s in self.species]) / AVOGADRO
        # reduced dipole moment delta* = mu^2 / (2 * 4 pi eps0 * eps * sigma^3)
        eps_j = eps * BOLTZMANN
        with np.errstate(divide="ignore", invalid="ignore"):
            delta = np.where(dip > 0, 0.5 * dip ** 2 / (4 * math.pi * EPS0 * eps_j * sig ** 3), 0.0)
        polar = dip > 0
        # pair parameters with the polar / non-polar correction (Hirschfelder, Curtiss & Bird)
        eps_jk = np.sqrt(np.outer(eps, eps))
        sig_jk = 0.5 * (sig[:, None] + sig[None, :])
        delta_jk = np.zeros((ns, ns))
        for j in range(ns):
            for k in range(ns):
                if polar[j] and polar[k]:
                    delta_jk[j, k] = 0.5 * dip[j] * dip[k] / (
                        4 * math.pi * EPS0 * eps_jk[j, k] * BOLTZMANN * sig_jk[j, k] ** 3)
                elif polar[j] != polar[k]:
                    p, n = (j, k) if polar[j] else (k, j)
                    alpha_n = alpha[n] / sig[n] ** 3
                    mu_p = dip[p] / math.sqrt(4 * math.pi * EPS0 * eps_j[p] * sig[p] ** 3)
                    xi = 1.0 + 0.25 * alpha_n * mu_p ** 2 * math.sqrt(eps[p] / eps[n])
                    eps_jk[j, k] *= xi * xi
                    sig_jk[j, k] *= xi ** (-1.0 / 6.0)
        visc = np.empty((T.size, ns))
        cond = np.empty((T.size, ns))
        for k in range(ns):
            ts = T / eps[k]
            om22 = _omega22(ts, delta[k])
            om11 = _omega11(ts, delta[k])
            mu = (5.0 / 16.0) * np.sqrt(math.pi * mass[k] * BOLTZMANN * T) / (math.pi * sig[k] ** 2 * om22)
            visc[:, k] = mu
            # self-diffusion (times pressure) for the internal-energy transport factor
            dp = (3.0 / 16.0) * np.sqrt(2.0 * math.pi / (mass[k] / 2.0)) * (BOLTZMANN * T) ** 1.5 / (
                math.pi * sig[k] ** 2 * om11)
            cp_r = self.cp_R(k, T)
            w = self.species[k].mw
            f_int = w / (R_UNIV * T) * dp / mu
            if self.species[k].geometry == 0:
                cv_rot = 0.0
            elif self.species[k].geometry == 1:
                cv_rot = 1.0
            else:
                cv_rot = 1.5
            a_fac = 2.5 - f_int
            fz = _parker(ts)
            fz298 = _parker(298.0 / eps[k])
            b_fac = self.species[k].rot_relax * fz298 / fz + 2.0 / math.pi * (5.0 / 3.0 * cv_rot + f_int)
            c1 = 2.0 / (math.pi * b_fac)
            cv_int = cp_r - 2.5 - cv_rot
            f_rot = f_int * (1.0 + c1 * a_fac)
            f_trans = 2.5 * (1.0 - c1 * cv_rot / 1.5 * a_fac)
            cond[:, k] = (mu / w) * R_UNIV * (f_trans * 1.5 + f_rot * cv_rot + f_int * cv_int)
        diff = np.empty((T.size, ns, ns))
        for j in range(ns):
            for k in range(ns):
                mjk = mass[j] * mass[k] / (mass[j] + mass[k])
                ts = T / eps_jk[j, k]
                om11 = _omega11(ts, delta_jk[j, k])
                diff[:, j, k] = (3.0 / 16.0) * np.sqrt(2.0 * math.pi / mjk) * (BOLTZMANN * T) ** 1.5 / (
                    p0 * math.pi * sig_jk[j, k] ** 2 * om11)
        return dict(visc_species=np.asfortranarray(visc), lambda_species=np.asfortranarray(cond),
                    diffbin_species=np.asfortranarray(diff))

    # ------------------------------------------------------------------ all
    def tables(self, p0: float = ONE_ATM, t_min_tab: float = 200.0, t_max_tab: float = 20000.0,
               dt_tab: float = 1.0, arr_preexp_multiplier: float = 1.0, transport: bool = True,
               kinetics: bool = True) -> "Tables":
        # num_t_tab = nint(t_max_tab-t_min_tab)/dt_tab ; dt_tab = (t_max-t_min)/num_t_tab
        # (multideal.F90:3133-3134): integer division of a nint() by a real -> real,
        # assigned to an integer (truncation).
        num_t_tab = int(float(int(round(t_max_tab - t_min_tab))) / dt_tab)
        dt_tab = (t_max_tab - t_min_tab) / num_t_tab
        t_tab = t_min_tab + np.arange(num_t_tab + 1) * dt_tab
        th = self.thermo_tables()
        out = dict(th)
        out.update(names=list(self.names), t_min_tab=t_min_tab, t_max_tab=t_max_tab, dt_tab=dt_tab,
                   num_t_tab=num_t_tab, p0=p0)
        if transport:
            out.update(self.transport_tables(t_tab, p0))
        if kinetics and self.n_reactions > 0:
            k = self.kinetics_tables(t_tab)
            k["arr_a"] = k["arr_a"] * arr_preexp_multiplier
            out.update(k)
        out["nreactions"] = self.n_reactions if kinetics else 0
        h298 = np.array([self.h_RT_s_R(k, np.array([298.15]))[0][0] * R_UNIV * 298.15
                         for k in range(self.n_species)])
        out["h298"] = h298
        # transport data in the units of fluid_prop.bin (tools/preproc_multi/preproc.py:136-140):
        # diameter m, well depth / k_B K, dipole C m, polarizability m^3
        out["sigma"] = np.array([s.diameter for s in self.species]) * 1e-10
        out["eps_k"] = np.array([s.well_depth for s in self.species])
        out["dipole"] = np.array([s.dipole for s in self.species]) * DEBYE
        out["polariz"] = np.array([s.polarizability for s in self.species]) * 1e-30
        out["geometry"] = np.array([s.geometry for s in self.species], dtype=np.int32)
        return Tables(**out)


def _omega22(ts, delta):
    """Reduced Ω(2,2)* — Neufeld, Janzen & Aziz (1972) with Brokaw's polar term."""
    om = 1.16145 * ts ** (-0.14874) + 0.52487 * np.exp(-0.77320 * ts) + 2.16178 * np.exp(-2.43787 * ts)
    return om + 0.2 * delta * delta / ts


def _omega11(ts, delta):
    """Reduced Ω(1,1)* — Neufeld, Janzen & Aziz (1972) with Brokaw's polar term."""
    om = (1.06036 * ts ** (-0.15610) + 0.19300 * np.exp(-0.47635 * ts)
          + 1.03587 * np.exp(-1.52996 * ts) + 1.76474 * np.exp(-3.89411 * ts))
    return om + 0.19 * delta * delta / ts


def _parker(ts):
    """Parker's rotational-relaxation temperature factor F(T*)."""
    x = 1.0 / np.asarray(ts, dtype=np.float64)
    return 1.0 + 0.5 * math.pi ** 1.5 * np.sqrt(x) + (0.25 * math.pi ** 2 + 2.0) * x + math.pi ** 1.5 * x ** 1.5


@dataclasses.dataclass
class Tables:
    """All tables the hot path consumes, in the reference's Fortran shapes."""
    names: List[str]
    indx_cp_l: int
    indx_cp_r: int
    nsetcv: int
    trange: np.ndarray
    cp_coeff: np.ndarray
    cv_coeff: np.ndarray
    rgas: np.ndarray
    mw: np.ndarray
    t_min_tab: float
    t_max_tab: float
    dt_tab: float
    num_t_tab: int
    p0: float
    h298: np.ndarray
    nreactions: int = 0
    visc_species: Optional[np.ndarray] = None
    lambda_species: Optional[np.ndarray] = None
    diffbin_species: Optional[np.ndarray] = None
    reac_ty: Optional[np.ndarray] = None
    isRev: Optional[np.ndarray] = None
    arr_a: Optional[np.ndarray] = None
    arr_b: Optional[np.ndarray] = None
    arr_ea: Optional[np.ndarray] = None
    falloff_coeffs: Optional[np.ndarray] = None
    tb_eff: Optional[np.ndarray] = None
    r_coeffs: Optional[np.ndarray] = None
    p_coeffs: Optional[np.ndarray] = None
    kc_tab: Optional[np.ndarray] = None
    # species transport data (fluid_prop.bin records 6-10; used by the reference only for
    # its Soret collision integrals)
    sigma: Optional[np.ndarray] = None
    eps_k: Optional[np.ndarray] = None
    dipole: Optional[np.ndarray] = None
    polariz: Optional[np.ndarray] = None
    geometry: Optional[np.ndarray] = None

    @property
    def n_species(self) -> int:
        return len(self.names)

    @property
    def mwinv(self) -> np.ndarray:
        return 1.0 / self.mw

    # ------------------------------------------------------------ host thermo
    def _nrange(self, T: float, k: int) -> int:
        # locate (multideal.F90:2633-2648), clamped to a valid range
        jl, ju = 0, self.nsetcv + 2
        while ju - jl > 1:
            jm = (ju + jl) // 2
            if T >= self.trange[k, jm - 1]:
                jl = jm
            else:
                ju = jm
        return min(max(jl, 1), self.nsetcv)

    def e_from_T(self, T: float, Y: Sequence[float]) -> float:
        """Mixture internal energy (get_e_from_temperature, multideal.F90:2622-2671)."""
        lo, hi = self.indx_cp_l, self.indx_cp_r
        nr = [self._nrange(T, k) for k in range(self.n_species)]
        ee = 0.0
        for k in range(self.n_species):
            ee += self.cv_coeff[hi + 1 - lo, k, nr[k] - 1] * Y[k]
        for l in range(lo, hi + 1):
            tprod = math.log(T) if l == -1 else T ** (l + 1) / (l + 1.0)
            cvl = 0.0
            for k in range(self.n_species):
                cvl += self.cv_coeff[l - lo, k, nr[k] - 1] * Y[k]
            ee += cvl * tprod
        return ee

    def e_from_T_many(self, T: np.ndarray, Y: np.ndarray) -> np.ndarray:
        """e_from_T for many states at once: T (n,), Y (n, N_S); the same sums in the same
        order, vectorised over the states."""
        lo, hi = self.indx_cp_l, self.indx_cp_r
        T = np.asarray(T, dtype=np.float64)
        Y = np.asarray(Y, dtype=np.float64)
        ns = self.n_species
        nr = np.ones((T.size, ns), dtype=np.int64)
        for m in range(2, self.nsetcv + 1):
            nr = np.where(T[:, None] >= self.trange[None, :, m - 1], m, nr)
        cols = np.arange(ns)[None, :]
        ee = np.zeros(T.size)
        for k in range(ns):
            ee = ee + self.cv_coeff[hi + 1 - lo, k, nr[:, k] - 1] * Y[:, k]
        for l in range(lo, hi + 1):
            tprod = np.log(T) if l == -1 else T ** (l + 1) / (l + 1.0)
            cvl = np.zeros(T.size)
            for k in range(ns):
                cvl = cvl + self.cv_coeff[l - lo, k, nr[:, k] - 1] * Y[:, k]
            ee = ee + cvl * tprod
        del cols
        return ee

    def cp_from_T_many(self, T: np.ndarray, Y: np.ndarray) -> np.ndarray:
        """cp_from_T for many states: T (n,), Y (n, N_S)."""
        lo, hi = self.indx_cp_l, self.indx_cp_r
        T = np.asarray(T, dtype=np.float64)
        Y = np.asarray(Y, dtype=np.float64)
        ns = self.n_species
        nr = np.ones((T.size, ns), dtype=np.int64)
        for m in range(2, self.nsetcv + 1):
            nr = np.where(T[:, None] >= self.trange[None, :, m - 1], m, nr)
        cp = np.zeros(T.size)
        for l in range(lo, hi + 1):
            cpl = np.zeros(T.size)
            for k in range(ns):
                cpl = cpl + self.cp_coeff[l - lo, k, nr[:, k] - 1] * Y[:, k]
            cp = cp + cpl * T ** l
        return cp

    def cp_from_T(self, T: float, Y: Sequence[float]) -> float:
        """Mixture c_p (get_cp, multideal.F90:5720-5761)."""
        lo, hi = self.indx_cp_l, self.indx_cp_r
        nr = [self._nrange(T, k) for k in range(self.n_species)]
        cp = 0.0
        for l in range(lo, hi + 1):
            cpl = 0.0
            for k in range(self.n_species):
                cpl += self.cp_coeff[l - lo, k, nr[k] - 1] * Y[k]
            cp += cpl * T ** l
        return cp

    def rmix(self, Y: Sequence[float]) -> float:
        r = 0.0
        for k in range(self.n_species):
            r += self.rgas[k] * Y[k]
        return r

    # ------------------------------------------------------- fluid_prop.bin I/O
    def write_fluid_prop(self, path: str):
        """Write the Fortran sequential-unformatted file read by set_fluid_prop
        (multideal.F90:2763-2926, written by tools/preproc_multi/preproc.py:196-246).

        Records carry 4-byte length markers.  Species names are character(len=30)
        (multideal.F90:213, preproc.py:197-198).  The transport tables are stored at
        1 atm, as the reader rescales them by 101325/p0.
        """
        ns = self.n_species
        z = np.zeros(ns)
        recs = []

        def rec(*arrs):
            payload = b"".join(a if isinstance(a, bytes) else np.asarray(a).tobytes(order="F") for a in arrs)
            recs.append(struct.pack("<i", len(payload)) + payload + struct.pack("<i", len(payload)))

        rec(np.array([ns, self.nreactions], dtype=np.int32))
        rec(b"".join(n.ljust(NAME_LEN)[:NAME_LEN].encode() for n in self.names))
        rec(self.mw.astype(np.float64))
        rec(self.rgas.astype(np.float64))
        rec(self.h298.astype(np.float64))
        for arr in (self.sigma, self.eps_k, self.dipole, self.polariz):
            rec((z if arr is None else np.asarray(arr)).astype(np.float64))
        rec((np.zeros(ns, dtype=np.int32) if self.geometry is None else np.asarray(self.geometry, dtype=np.int32)))
        ncoef = self.indx_cp_r + 3 - self.indx_cp_l
        rec(np.array([ncoef, self.nsetcv], dtype=np.int32))
        rec(np.asfortranarray(self.trange))
        rec(np.asfortranarray(self.cp_coeff / self.rgas[None, :, None]))
        rec(np.array([self.t_min_tab, self.t_max_tab, self.dt_tab], dtype=np.float64))
        rec(np.asfortranarray(self.visc_species))
        rec(np.asfortranarray(self.lambda_species))
        rec(np.asfortranarray(self.diffbin_species * self.p0 / ONE_ATM))
        if self.nreactions > 0:
            rec(self.reac_ty.astype(np.int32))
            for arr in (self.arr_a, self.arr_b, self.arr_ea, self.falloff_coeffs, self.tb_eff,
                        self.r_coeffs, self.p_coeffs, self.kc_tab):
                rec(np.asfortranarray(arr))
        with open(path, "wb") as fh:
            fh.write(b"".join(recs))


def read_fluid_prop(path: str, p0: float) -> Tables:
    """Read a fluid_prop.bin (set_fluid_prop, multideal.F90:2763-2926)."""
    with open(path, "rb") as fh:
        raw = fh.read()
    recs = []
    off = 0
    while off < len(raw):
        (n,) = struct.unpack_from("<i", raw, off)
        recs.append(raw[off + 4: off + 4 + n])
        off += 8 + n
    it = iter(recs)
    ns, nr = np.frombuffer(next(it), dtype=np.int32)
    ns, nr = int(ns), int(nr)
    nraw = next(it)
    if len(nraw) != NAME_LEN * ns:
        raise ValueError(f"species-name record of {len(nraw)} bytes; set_fluid_prop reads {ns} x {NAME_LEN}")
    names = [nraw[i * NAME_LEN:(i + 1) * NAME_LEN].decode().strip() for i in range(ns)]
    mw = np.frombuffer(next(it), dtype=np.float64).copy()
    rgas = np.frombuffer(next(it), dtype=np.float64).copy()
    h298 = np.frombuffer(next(it), dtype=np.float64).copy()
    tr = {k: np.frombuffer(next(it), dtype=np.float64).copy() for k in ("sigma", "eps_k", "dipole", "polariz")}
    tr["geometry"] = np.frombuffer(next(it), dtype=np.int32).copy()
    ncoef, nsetcv = (int(v) for v in np.frombuffer(next(it), dtype=np.int32))
    lo, hi = {7: (0, 4), 9: (-2, 4)}[ncoef]
    trange = np.frombuffer(next(it), dtype=np.float64).reshape((ns, nsetcv + 1), order="F").copy(order="F")
    cp = np.frombuffer(next(it), dtype=np.float64).reshape((ncoef, ns, nsetcv), order="F").copy(order="F")
    cp = cp * rgas[None, :, None]
    cv = cp.copy(order="F")
    cv[0 - lo] = cp[0 - lo] - rgas[:, None]
    tmin, tmax, dtt = np.frombuffer(next(it), dtype=np.float64)
    num = int(float(int(round(tmax - tmin))) / dtt)
    dtt = (tmax - tmin) / num
    visc = np.frombuffer(next(it), dtype=np.float64).reshape((num + 1, ns), order="F").copy(order="F")
    lam = np.frombuffer(next(it), dtype=np.float64).reshape((num + 1, ns), order="F").copy(order="F")
    dif = np.frombuffer(next(it), dtype=np.float64).reshape((num + 1, ns, ns), order="F").copy(order="F")
    dif = dif * ONE_ATM / p0
    kw = {}
    if nr > 0:
        kw["reac_ty"] = np.frombuffer(next(it), dtype=np.int32).copy()
        # the file carries no reversibility flags (set_fluid_prop allocates isRev without
        # reading it); every reaction of the reference's example inputs is reversible
        kw["isRev"] = np.ones(nr, dtype=np.int32)
        shapes = [(nr, 2), (nr, 2), (nr, 2), (nr, 5), (nr, ns), (nr, ns), (nr, ns), (num + 1, nr)]
        keys = ["arr_a", "arr_b", "arr_ea", "falloff_coeffs", "tb_eff", "r_coeffs", "p_coeffs", "kc_tab"]
        for key, shp in zip(keys, shapes):
            kw[key] = np.frombuffer(next(it), dtype=np.float64).reshape(shp, order="F").copy(order="F")
        kc = kw["kc_tab"]
        kw["kc_tab"] = np.clip(kc, np.finfo(np.float64).tiny, np.finfo(np.float64).max)
    return Tables(names=names, indx_cp_l=lo, indx_cp_r=hi, nsetcv=nsetcv, trange=trange, cp_coeff=cp,
                  cv_coeff=cv, rgas=rgas, mw=mw, t_min_tab=float(tmin), t_max_tab=float(tmax),
                  dt_tab=float(dtt), num_t_tab=num, p0=p0, h298=h298, nreactions=nr, visc_species=visc,
                  lambda_species=lam, diffbin_species=dif, **tr, **kw)
