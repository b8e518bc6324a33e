"""Seeded synthetic inputs for the measured configurations.

* ``h2_air_mechanism`` — the 9-species H2/O2/N2 set: the laminar-premixed-flame
  mechanism of the reference's examples (11 species, 21 reactions incl. three-body and
  Troe falloff) with Ar and He removed (they appear there only as collision partners).
* ``turbulence_box`` — periodic reacting decaying-turbulence cube: lean H2-air (phi=0.5,
  X_H2:X_O2:X_N2 = 1:1:3.76), uniform T and p, solenoidal velocity from seeded random
  phases with a Passot-Pouquet spectrum E(k) ~ (k/k0)^4 exp(-2 (k/k0)^2), scaled to a
  turbulent Mach number M_t = sqrt(<u.u>)/c.  Domain side 2*pi mm.
* ``reactor_batch`` — independent 0-D reactors: T log-uniform, p log-uniform, Y from a
  seeded Dirichlet(1) over the species, rho from the equation of state, at rest.
* ``shock_tube`` — the reactive shock-tube initial state of init_reactive_tube
  (multideal.F90:5305-5381) extruded along x.
* ``mixing_layer`` — syngas / air temporal mixing layers on a periodic box (the 14-species
  Keromnes-2013 mechanism): a fuel band (H2:CO = 1:1 molar, 300 K) between two tanh
  interfaces in y, hot air (1500 K) outside, velocity jump from a convective Mach number
  M_c = dU / (c_fuel + c_air), uniform pressure, and a seeded 1 % broadband (solenoidal)
  velocity perturbation.  Two interfaces make the y direction periodic, so the one-rank box
  needs no inflow or outflow treatment.

Arrays are returned as ``w[m, k, j, i]`` (C order, i.e. the reference's Fortran layout
``w(i, j, k, m)``) including the ng ghost layers, which are filled with the periodic
extension for periodic axes.
"""
from __future__ import annotations

import math
import pathlib
from typing import Optional, Sequence

import numpy as np

from .mechanism import Mechanism, R_UNIV, Tables

MECH_DIR = pathlib.Path(__file__).resolve().parent / "mechanisms"
H2_AIR_9 = ["H2", "H", "O2", "O", "OH", "HO2", "H2O2", "H2O", "N2"]


def mechanism_path(name: str) -> pathlib.Path:
    return MECH_DIR / name


def h2_air_mechanism() -> Mechanism:
    return Mechanism(str(mechanism_path("h2_premix_11sp.yaml")), species=H2_AIR_9)


def mole_to_mass(tables: Tables, x: dict) -> np.ndarray:
    xs = np.array([x.get(n, 0.0) for n in tables.names])
    y = xs * tables.mw
    return y / y.sum()


def _pad_periodic(a: np.ndarray, ng: int) -> np.ndarray:
    return np.pad(a, [(0, 0)] + [(ng, ng)] * (a.ndim - 1), mode="wrap")


def solenoidal_field(n, seed: int, k0: float = 4.0) -> np.ndarray:
    """Random-phase solenoidal velocity (3, nz, ny, nx) with unit rms speed sqrt(<u.u>)=1,
    periodic on the box; n is the cube edge or the shape (nz, ny, nx)."""
    nz, ny, nx = (n, n, n) if np.isscalar(n) else tuple(n)
    rng = np.random.default_rng(seed)
    kz1 = np.fft.fftfreq(nz, d=1.0 / nz)
    ky1 = np.fft.fftfreq(ny, d=1.0 / ny)
    kr = np.fft.rfftfreq(nx, d=1.0 / nx)
    kz, ky, kx = np.meshgrid(kz1, ky1, kr, indexing="ij")
    kk = np.sqrt(kx * kx + ky * ky + kz * kz)
    kk[0, 0, 0] = 1.0
    spec = (kk / k0) ** 4 * np.exp(-2.0 * (kk / k0) ** 2)
    amp = np.sqrt(spec / (4.0 * math.pi * kk * kk))
    amp[0, 0, 0] = 0.0
    shape = kk.shape
    xi = [(rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) * amp for _ in range(3)]
    # project onto the plane normal to k
    kdot = (kx * xi[0] + ky * xi[1] + kz * xi[2]) / (kk * kk)
    uh = [xi[0] - kx * kdot, xi[1] - ky * kdot, xi[2] - kz * kdot]
    u = np.stack([np.fft.irfftn(c, s=(nz, ny, nx), axes=(0, 1, 2)) for c in uh])    # axes (k, j, i)
    u -= u.mean(axis=(1, 2, 3), keepdims=True)
    urms = math.sqrt(float((u * u).sum(axis=0).mean()))
    return u / urms


def turbulence_box(tables: Tables, n: int, seed: int = 2024, ng: int = 3, temperature: float = 1200.0,
                   pressure: float = 101325.0, mach_t: float = 0.3, k0: float = 4.0,
                   x: Optional[dict] = None):
    """Conserved variables (nv, n+2ng, n+2ng, n+2ng) and the temperature guess."""
    x = x or {"H2": 1.0, "O2": 1.0, "N2": 3.76}
    y = mole_to_mass(tables, x)
    rmix = tables.rmix(y)
    rho = pressure / (rmix * temperature)
    cp = tables.cp_from_T(temperature, y)
    c = math.sqrt(cp / (cp - rmix) * rmix * temperature)
    vel = solenoidal_field(n, seed, k0) * (mach_t * c)
    e = tables.e_from_T(temperature, y)
    ns = tables.n_species
    w = np.empty((ns + 4, n, n, n))
    for s in range(ns):
        w[s] = rho * y[s]
    w[ns + 0] = rho * vel[0]      # vel axes are (k, j, i): component order x, y, z
    w[ns + 1] = rho * vel[1]
    w[ns + 2] = rho * vel[2]
    w[ns + 3] = rho * (e + 0.5 * (vel * vel).sum(axis=0))
    wp = _pad_periodic(w, ng)
    tg = np.full(wp.shape[1:], temperature)
    info = dict(rho0=rho, t0=temperature, p0=pressure, c0=c, u_rms=mach_t * c, length=2 * math.pi * 1e-3, y=y)
    return wp, tg, info


def reactor_batch(tables: Tables, n: int, seed: int = 7, t_range=(800.0, 2500.0),
                  p_range=(0.5 * 101325.0, 20 * 101325.0), alpha: float = 1.0):
    """n independent reactors at rest: SoA conserved state w[nv, n] and the T guess; T and p
    log-uniform on the given ranges, Y from a seeded Dirichlet(alpha) over the species."""
    rng = np.random.default_rng(seed)
    ns = tables.n_species
    T = np.exp(rng.uniform(math.log(t_range[0]), math.log(t_range[1]), n))
    p = np.exp(rng.uniform(math.log(p_range[0]), math.log(p_range[1]), n))
    Y = rng.dirichlet(np.full(ns, alpha), size=n)
    rmix = Y @ tables.rgas
    rho = p / (rmix * T)
    w = np.zeros((ns + 4, n))
    w[:ns] = (rho[:, None] * Y).T
    w[ns + 3] = rho * tables.e_from_T_many(T, Y)
    return w, T.copy(), dict(T=T, p=p, Y=Y)


def shock_tube(tables: Tables, nx: int, init_y: Sequence[float], rho0: float, p0: float,
               ng: int = 3, length: float = 0.12, ny: int = 1, nz: int = 1):
    """init_reactive_tube: quiescent (rho0, p0) for x < x(nx)/2, driven state beyond."""
    ns = tables.n_species
    y = np.asarray(init_y, dtype=np.float64)
    rmix = tables.rmix(y)
    t0 = p0 / rho0 / rmix
    e0 = tables.e_from_T(t0, y)
    win = np.zeros(ns + 4)
    win[:ns] = rho0 * y
    win[ns + 3] = rho0 * e0
    rho_t, u_t, p_t = 0.18075, -487.34, 35594.0
    t_t = p_t / rmix / rho_t
    e_t = tables.e_from_T(t_t, y)
    wt = np.zeros(ns + 4)
    wt[:ns] = rho_t * y
    wt[ns] = rho_t * u_t
    wt[ns + 3] = rho_t * (e_t + 0.5 * u_t * u_t)
    dx = length / (nx - 1)
    xs = (np.arange(1, nx + 1) - 1) * dx
    w = np.zeros((ns + 4, nz, ny, nx))
    left = xs < 0.5 * xs[-1]
    w[:, :, :, left] = win[:, None, None, None]
    w[:, :, :, ~left] = wt[:, None, None, None]
    wp = np.pad(w, [(0, 0), (ng, ng), (ng, ng), (ng, ng)], mode="edge")
    tg = np.where(np.pad(np.broadcast_to(left, (nz, ny, nx)), [(ng, ng)] * 3, mode="edge"), t0, t_t)
    return wp, tg, dict(t0=t0, t_tube=t_t, rho0=rho0, p0=p0)


def flow_params(tables: Tables, y: Sequence[float], rho0: float, p0: float, u0: float = 0.0) -> dict:
    """Reference quantities of set_flow_params (multideal.F90:3415-3539): T0 from the
    equation of state, the sound speed c0, and u0 replaced by c0 when it is below the Newton
    tolerance (u0 scales the WENO smoothness indicators and sets eps_sensor = u0^2)."""
    y = np.asarray(y, dtype=np.float64)
    rmix = tables.rmix(y)
    t0 = p0 / (rho0 * rmix)
    cp = tables.cp_from_T(t0, y)
    c0 = math.sqrt(cp / (cp - rmix) * rmix * t0)
    u = c0 if abs(u0) < 1e-9 else u0
    return dict(t0=t0, c0=c0, u0=u, rho0=rho0, p0=p0, eps_sensor=u * u)


def reactor_state(tables: Tables, shape: Sequence[int], ng: int, t0: float, p0: float, y: Sequence[float]):
    """init_reactor (multideal.F90:5608-5669): every cell at rest at (t0, p0, Y); returns the
    padded conserved variables (nv, nz+2ng, ny+2ng, nx+2ng) and the temperature guess."""
    nx, ny, nz = shape
    y = np.asarray(y, dtype=np.float64)
    rho = p0 / tables.rmix(y) / t0
    e = tables.e_from_T(t0, y)
    ns = tables.n_species
    w = np.zeros((ns + 4, nz + 2 * ng, ny + 2 * ng, nx + 2 * ng))
    w[:ns] = (rho * y)[:, None, None, None]
    w[ns + 3] = rho * e
    return w, np.full(w.shape[1:], t0)


def shock_tube_mechanism() -> Mechanism:
    """The reactive shock-tube input (9 species H2/O2/.../Ar, 18 reactions)."""
    return Mechanism(str(mechanism_path("h2_shocktube_9sp.yaml")))


# init_Y of the reactive shock-tube example: H2:O2:Ar = 2:1:7 molar
SHOCK_TUBE_Y = {"H2": 0.01267282566, "O2": 0.10138260528, "Ar": 0.88594456906}


def keromnes_mechanism() -> Mechanism:
    return Mechanism(str(mechanism_path("keromnes2013_14sp.yaml")))


def mixing_layer(tables: Tables, nx: int, ny: int, nz: int, seed: int = 2024, ng: int = 3, dx: float = 1.0e-5,
                 mach_c: float = 0.8, t_fuel: float = 300.0, t_air: float = 1500.0, pressure: float = 101325.0,
                 thickness: float = 1.0 / 40.0, perturbation: float = 0.01):
    """Conserved variables (nv, nz+2ng, ny+2ng, nx+2ng), the temperature guess and the box
    lengths of the double temporal mixing layer; the band's fuel fraction is
    phi(y) = (tanh((y - Ly/4)/d) - tanh((y - 3Ly/4)/d)) / 2 with d = thickness * Ly, and Y and
    T are the phi-weighted means of the two streams."""
    y_fuel = mole_to_mass(tables, {"H2": 1.0, "CO": 1.0})
    y_air = mole_to_mass(tables, {"O2": 1.0, "N2": 3.76})

    def sound(y, t):
        r = tables.rmix(y)
        cp = tables.cp_from_T(t, y)
        return math.sqrt(cp / (cp - r) * r * t)

    du = mach_c * (sound(y_fuel, t_fuel) + sound(y_air, t_air))
    ly = ny * dx
    yc = (np.arange(ny) + 0.5) * dx
    d = thickness * ly
    phi = 0.5 * (np.tanh((yc - 0.25 * ly) / d) - np.tanh((yc - 0.75 * ly) / d))
    ns = tables.n_species
    Y = phi[:, None] * y_fuel[None, :] + (1.0 - phi[:, None]) * y_air[None, :]          # (ny, ns)
    T = phi * t_fuel + (1.0 - phi) * t_air
    rho = pressure / ((Y @ tables.rgas) * T)
    e = tables.e_from_T_many(T, Y)
    u = du * (phi - 0.5)
    # broadband perturbation: a solenoidal field on the cube that holds the box, cropped
    pert = solenoidal_field((nz, ny, nx), seed) * (perturbation * du)
    vel = pert.copy()
    vel[0] += u[None, :, None]
    w = np.empty((ns + 4, nz, ny, nx))
    for s in range(ns):
        w[s] = (rho * Y[:, s])[None, :, None]
    rb = rho[None, :, None]
    w[ns + 0] = rb * vel[0]
    w[ns + 1] = rb * vel[1]
    w[ns + 2] = rb * vel[2]
    w[ns + 3] = rb * (e[None, :, None] + 0.5 * (vel * vel).sum(axis=0))
    wp = _pad_periodic(w, ng)
    tg = np.pad(np.broadcast_to(T[None, :, None], (nz, ny, nx)), [(ng, ng)] * 3, mode="wrap")
    info = dict(rho0=float(rho.mean()), u0=du, length=[nx * dx, ny * dx, nz * dx], du=du)
    return wp, np.ascontiguousarray(tg), info


def weno_threshold_for_fraction(ducros: np.ndarray, ws: int, fraction: float = 0.10) -> float:
    """Sensor threshold at which about `fraction` of the x faces take WENO: face i+1/2 is
    shocked when a cell of its 2*ws-wide stencil has Ducros above the threshold
    (shock sensing of the euler kernels).  ducros: interior (nz, ny, nx), periodic in x."""
    win = np.max(np.stack([np.roll(ducros, -s, axis=2) for s in range(-ws + 1, ws + 1)]), axis=0)
    q = float(np.quantile(win, 1.0 - fraction))
    # midway between two distinct sensor values, so that no cell sits on the threshold (a
    # cell exactly at it would flip with the last bit of its sensor value)
    u = np.unique(win)
    i = int(np.searchsorted(u, q))
    if 0 < i < u.size:
        return float(0.5 * (u[i - 1] + u[i]))
    return q

