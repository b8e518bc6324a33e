"""Point-wise auxiliaries on the oracle: a list of cell states laid out as an nx*1*1 box
(test infrastructure)."""
import numpy as np


def point_aux(pkg, backend, tables, w_points, t_guess=1000.0, p0=None):
    """eval_aux of each column of w_points (nv, npts); returns w_aux (nv_aux, npts)."""
    ns = tables.n_species
    npts = w_points.shape[1]
    ng = 3
    num = pkg.solver.Numerics(enable_chemistry=0)
    s = pkg.solver.Solver(backend, tables, npts, 1, 1, ng, [1.0, 1.0, 1.0], numerics=num,
                          p0=tables.p0 if p0 is None else p0)
    px = npts + 2 * ng
    w = np.empty((ns + 4, 2 * ng + 1, 2 * ng + 1, px))
    cols = np.concatenate([np.repeat(w_points[:, :1], ng, axis=1), w_points,
                           np.repeat(w_points[:, -1:], ng, axis=1)], axis=1)
    w[:] = cols[:, None, None, :]
    s.set_state(w, np.full(w.shape[1:], float(t_guess)))
    s.eval_aux()
    wa = backend.to_numpy(s.w_aux)
    return wa[:, ng, ng, ng:ng + npts]


def quiescent(tables, temperature, pressure, y):
    """Conserved variables of quiescent states; temperature/pressure arrays (npts), y (ns, npts)."""
    ns = tables.n_species
    temperature = np.atleast_1d(np.asarray(temperature, dtype=np.float64))
    y = np.asarray(y, dtype=np.float64).reshape(ns, -1)
    npts = max(y.shape[1], temperature.size)
    temperature = np.broadcast_to(temperature, (npts,))
    pressure = np.broadcast_to(np.asarray(pressure, dtype=np.float64), (npts,))
    y = np.broadcast_to(y, (ns, npts))
    w = np.zeros((ns + 4, npts))
    for b in range(npts):
        rho = pressure[b] / (tables.rmix(y[:, b]) * temperature[b])
        w[:ns, b] = rho * y[:, b]
        w[ns + 3, b] = rho * tables.e_from_T(temperature[b], y[:, b])
    return w
