"""Driver-side steps either side of the hot path (host, numpy/scipy):
seeded velocity models for the benchmark configurations, the Ricker
wavelet, edge padding, migration-model smoothing, acquisition geometry and
the 1-based padded index convention.

Reference anchors:
  Ricker            SRC/modeling/main_Acoustic_Data.m:134-136
  dx/dt rule        SRC/modeling/main_Acoustic_Data.m:66-86
  edge padding      SRC/modeling/main_Acoustic_Data.m:144-145
  index convention  SRC/modeling/main_Acoustic_Data.m:120-126 (xs = d_pml + x/dx)
  geometry          SRC/MNAD/acquisi.m:47-169
  smoothing         SRC/MNAD/v_smooth.m (box filter on slowness); the
                    benchmark configurations use a Gaussian instead.
"""
from __future__ import annotations

import numpy as np


def ricker(fc: float, dt: float, nt: int) -> np.ndarray:
    """w = (1 - 2 tau^2) exp(-tau^2), tau = pi (t - 1.1/fc) fc, t = (0:nt-1) dt."""
    t = np.arange(int(nt)) * float(dt)
    tau = np.pi * (t - 1.1 / fc) * fc
    return (1 - 2 * tau**2) * np.exp(-tau**2)


def reference_dx_dt(vmin: float, vmax: float, fc: float, T_max: float, with_end: bool = False):
    """main_Acoustic_Data.m:66-86: dx = vmin/(2.5*3fc); dt = 0.8*0.65*dx/vmax
    computed from the un-floored dx; then dx = floor(dx); nt = length(0:dt:T_max-dt).
    with_end: the example drivers' time axis t = 0:dt:T_max
    (examples/modeling/main_Acoustic_Data.m:77-79), nt = length(0:dt:T_max)."""
    dx = vmin / (2.5 * 3 * fc)
    dt = 0.8 * ((0.65 * dx) / vmax)
    last = T_max if with_end else T_max - dt
    nt = int(np.floor(last / dt + 1e-10)) + 1
    return float(np.floor(dx)), dt, nt


def pad_edges(V: np.ndarray, npml: int) -> np.ndarray:
    """Replicate edge columns then edge rows npml times (main_Acoustic_Data.m:144-145)."""
    return np.pad(np.asarray(V, dtype=np.float64), int(npml), mode="edge")


def v_smooth(V: np.ndarray, s1: int, s2: int, nit: int) -> np.ndarray:
    """v_smooth.m: nit passes of an s1 x s2 box filter on slowness with edge
    replication."""
    from scipy.ndimage import uniform_filter
    s = 1.0 / np.asarray(V, dtype=np.float64)
    for _ in range(int(nit)):
        s = uniform_filter(s, size=(int(s1), int(s2)), mode="nearest")
    return 1.0 / s


def gaussian_smooth(V: np.ndarray, sigma_m: float, dx: float, dz: float | None = None):
    """Gaussian smoothing of the velocity with a standard deviation of
    sigma_m metres (edge replication)."""
    from scipy.ndimage import gaussian_filter
    dz = dx if dz is None else dz
    return gaussian_filter(np.asarray(V, dtype=np.float64), sigma=(sigma_m / dz, sigma_m / dx),
                           mode="nearest")


def correlated_noise(shape, corr_m: float, dx: float, rng) -> np.ndarray:
    """Unit-variance Gaussian random field with a Gaussian correlation of
    length corr_m (white noise filtered with a Gaussian of sigma corr_m/sqrt 2)."""
    from scipy.ndimage import gaussian_filter
    g = gaussian_filter(rng.standard_normal(shape), sigma=corr_m / np.sqrt(2) / dx, mode="wrap")
    return g / g.std()


def model_c2(seed: int = 2024, nz: int = 501, nx: int = 1001, dx: float = 10.0):
    """Seeded smooth layered model 1500-4000 m/s with 5 % Gaussian-correlated
    perturbations (correlation length 50 m); migration model = the true model
    smoothed with a 100 m Gaussian.  Returns (V_true, V_migration), physical grid."""
    rng = np.random.default_rng(seed)
    z = np.arange(nz) * dx
    zmax = z[-1]
    base = 1500.0 + 2500.0 * (z / zmax) ** 1.2
    # a few smooth interfaces with seeded small undulations
    x = np.arange(nx) * dx
    V = np.tile(base[:, None], (1, nx))
    for k in range(4):
        z0 = zmax * (0.2 + 0.18 * k) + 40.0 * np.sin(2 * np.pi * x / (3000.0 + 700 * k) + rng.uniform(0, 6))
        V += np.where(z[:, None] > z0[None, :], 120.0, 0.0)
    V *= 1.0 + 0.05 * correlated_noise((nz, nx), 50.0, dx, rng)
    V = np.clip(V, 1500.0, 4000.0)
    return V, gaussian_smooth(V, 100.0, dx)


def model_c3(seed: int = 3, nz: int = 751, nx: int = 3401, dx: float = 10.0):
    """Seeded synthetic: 200 m water layer (1500 m/s), dipping layers
    1500-4000 m/s, three normal faults (throw 30-80 m) and an irregular
    4500 m/s body.  Returns (V_true, V_migration)."""
    rng = np.random.default_rng(seed)
    z = (np.arange(nz) * dx)[:, None]
    x = (np.arange(nx) * dx)[None, :]
    zmax, xmax = z[-1, 0], x[0, -1]
    dip = np.tan(np.deg2rad(4.0))
    depth = z - dip * (x - xmax / 2)
    # fault offsets: throw added to the depth coordinate right of each fault
    for xf, throw in zip(rng.uniform(0.2, 0.8, 3) * xmax, rng.uniform(30, 80, 3)):
        xfz = xf + 0.3 * z                     # fault plane dips
        depth = depth - np.where(x > xfz, throw, 0.0)
    nlay = 9
    bounds = np.linspace(200.0, zmax, nlay + 1)
    vel = np.linspace(1700.0, 4000.0, nlay)
    V = np.full((nz, nx), vel[0])
    for k in range(1, nlay):
        V = np.where(depth > bounds[k], vel[k], V)
    # irregular high-velocity body
    cx, cz = 0.6 * xmax, 0.55 * zmax
    ang = np.arctan2(z - cz, x - cx)
    rad = 600.0 * (1 + 0.25 * np.sin(3 * ang + rng.uniform(0, 6)) + 0.1 * np.cos(5 * ang))
    V = np.where(np.hypot(x - cx, (z - cz) * 1.6) < rad, 4500.0, V)
    V = np.where(z < 200.0, 1500.0, V)
    return V, gaussian_smooth(V, 100.0, dx)


def model_c5(seed: int = 5, nz: int = 2001, nx: int = 8001, dx: float = 12.5):
    """Large long-recording model: the faulted layered + high-velocity-body
    synthetic of model_c3 on an 8001 x 2001 grid at 12.5 m."""
    return model_c3(seed=seed, nz=nz, nx=nx, dx=dx)


def model_layered_random(rng, nz, nx, dx, vmin=1500.0, vmax=4500.0):
    """Layered-plus-random model of the serving workload (1500-4500 m/s)."""
    z = np.arange(nz) * dx
    base = vmin + (vmax - vmin) * 0.9 * (z / max(z[-1], 1.0))
    V = np.tile(base[:, None], (1, nx))
    V *= 1.0 + 0.03 * correlated_noise((nz, nx), 50.0, dx, rng)
    return np.clip(V, vmin, vmax)


def acquisi(V, dx, dz, zsc, zrc, h_offset, ds, dr, slag, rlag, acqType="fixspread",
            manualGeom=None):
    """acquisi.m:1-169: Source [2 x ns] and Receiver list of (x, z) metre
    arrays per shot, for the geometry types of the reference."""
    nz, nx = np.shape(V)
    ds = int(np.fix(ds / dx)) or 1
    dr = int(np.fix(dr / dx)) or 1
    kind = acqType.lower()
    if kind == "arbitrary":
        if manualGeom is None:
            raise ValueError("acquisi:missingManualGeom: acqType is 'arbitrary' but manualGeom was not provided.")
        for key in ("xs", "zs", "xr", "zr"):
            if key not in manualGeom:
                raise ValueError(f"acquisi:missingManualField: manualGeom is missing required field '{key}'.")
        xs = np.asarray(manualGeom["xs"], dtype=np.float64)
        zs = np.asarray(manualGeom["zs"], dtype=np.float64)
        xr = [np.asarray(a, dtype=np.float64) for a in manualGeom["xr"]]
        zr = [np.asarray(a, dtype=np.float64) for a in manualGeom["zr"]]
    elif kind in ("fixspread", "onesided", "splitspread", "obn"):
        colon = lambda a, s, b: np.arange(a, b + 1, s, dtype=np.float64) if b >= a else np.zeros(0)
        xs = colon(int(np.ceil(slag / dx)), ds, nx - int(np.floor(slag / dx)))
        zs = np.zeros_like(xs) + np.floor(zsc / dz)
        ns = xs.size
        h_off = int(np.floor(h_offset / dx))
        rend = nx - int(np.floor(rlag / dx))
        xr, zr = [], []
        if kind == "onesided":
            for s in xs:
                xr.append(colon(s, dr, rend) if s + h_off > rend else colon(s, dr, s + h_off - 1))
        elif kind == "splitspread":
            for s in xs:
                if s + h_off > rend:
                    xr.append(colon(s - h_off, dr, rend))
                elif s - h_off <= 0:
                    xr.append(colon(1, dr, s + h_off - 1))
                else:
                    xr.append(colon(s - h_off, dr, s + h_off - 1))
        elif kind == "obn":
            Vw = np.asarray(V)[0, 0]
            xs = colon(int(np.ceil(rlag / dx)), dr, rend)
            ns = xs.size
            zs = np.array([np.max(np.nonzero(np.asarray(V)[:, int(c) - 1] == Vw)[0]) + 1
                           for c in xs], dtype=np.float64)
            for _ in range(ns):
                xr.append(colon(int(np.ceil(slag / dx)), ds, nx - int(np.floor(slag / dx))))
        else:
            for _ in range(ns):
                xr.append(colon(int(np.ceil(rlag / dx)), dr, rend))
        if kind == "obn":
            zr = [0 * a + np.floor(zsc / dz) for a in xr]
        else:
            zr = [0 * a + np.floor(zrc / dz) for a in xr]
    else:
        raise ValueError(f"acquisi:unknownType: Unknown acqType \"{acqType}\".")
    Source = np.vstack([dx * xs, dz * zs])
    Receiver = [(dx * a, dz * b) for a, b in zip(xr, zr)]
    return Source, Receiver


def to_padded_indices(Source, Receiver, dx, dz, d_pml):
    """main_Acoustic_Data.m:120-126: 1-based padded indices d_pml + metres/dx."""
    xs = d_pml + Source[0] / dx
    zs = d_pml + Source[1] / dz
    xr = [d_pml + r[0] / dx for r in Receiver]
    zr = [d_pml + r[1] / dz for r in Receiver]
    return xs, zs, xr, zr
