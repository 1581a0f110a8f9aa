"""Synthetic meshes, models and (omega, ky) batches for the fem25 workloads.

Units follow the reference (mesh_utils.f:1120-1141): x, z in metres with z
positive up, velocities in m/s, density in kg/m^3; elastic coefficients are
stored per anchor node as ECOEFF(:,1) = lambda = rho (Vp^2 - 2 Vs^2) and
ECOEFF(:,2) = mu = rho Vs^2 (mesh_utils.f:1135-1138).

Meshes are structured bilinear quadrilaterals.  Anchors are numbered
row-major from the bottom-left corner and each element lists its anchors
counter-clockwise (bottom-left, bottom-right, top-right, top-left), the
pattern of GEN_QMESH_LOCS (disp25.f90:345-412).  Element e = iz*nx + ix.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np


@dataclasses.dataclass
class Mesh:
    nord: int
    nx: int
    nz: int
    ieng: np.ndarray      # int32 (nelem, 4), 1-based anchor numbers (IENG(1:4, e))
    xlocs: np.ndarray     # float64 (nnpg,)
    zlocs: np.ndarray
    cdomain: bytes        # one char per element: b'I', b'E' or b'A'
    cnnpg: bytes          # two chars per anchor
    xd: np.ndarray        # distance into the PML in x at anchors (m)
    zd: np.ndarray
    xwidth: float = 0.0
    zwidth: float = 0.0

    @property
    def nelem(self) -> int:
        return self.ieng.shape[0]

    @property
    def nnpg(self) -> int:
        return self.xlocs.shape[0]

    @property
    def nl(self) -> int:
        return self.nord + 1

    @property
    def nen(self) -> int:
        return self.nl * self.nl


@dataclasses.dataclass
class Model:
    dens: np.ndarray      # (nnpg,) kg/m^3
    ecoeff: np.ndarray    # (nnpg, 2): lambda, mu (Pa), Fortran ECOEFF(NNPG,2)
    vp: np.ndarray
    vs: np.ndarray
    # dispersive medium (LDISP): quality factors at the anchors and the
    # reference frequency FREQ0 (Hz) of VCMPLX (haskattn.f:602-641)
    qp: np.ndarray | None = None
    qs: np.ndarray | None = None
    freq0: float = 0.0


def with_attenuation(model: Model, qp: float | np.ndarray, qs: float | np.ndarray,
                     freq0: float) -> Model:
    """The same elastic model with quality factors QP, QS at the anchors."""
    n = model.dens.size
    return dataclasses.replace(model, qp=np.broadcast_to(np.asarray(qp, np.float64), (n,)).copy(),
                               qs=np.broadcast_to(np.asarray(qs, np.float64), (n,)).copy(),
                               freq0=float(freq0))


@dataclasses.dataclass
class PML:
    """PML parameters as selected by ASMBLE_DRIVER's CFTYPE (asmble.f90:78-88)."""
    rcoeff: float = 1.0e-4
    dmax: float = 0.0
    kappa: float = 1.0
    fcent: float = 0.0

    def as_array(self, mesh: Mesh) -> np.ndarray:
        return np.array([mesh.xwidth, mesh.zwidth, self.rcoeff, self.dmax, self.kappa,
                         self.fcent], dtype=np.float64)


def structured_mesh(nx: int, nz: int, nord: int, width: float = 80e3, depth: float = 40e3,
                    npml: int = 0, nbielak: int = 0) -> Mesh:
    """Structured nx*nz mesh over x in [0, width], z in [-depth, 0].

    With npml > 0 the outer npml element layers on the left, right and bottom
    are absorbing ('A'); the next nbielak layers are the Bielak ring ('E');
    the rest is interior ('I').  Outer anchors of the PML are fixed ('FL',
    'FR', 'FB'), top anchors are free surface ('FS').
    """
    dx = width / nx
    dz = depth / nz
    ix = np.arange(nx + 1)
    iz = np.arange(nz + 1)
    X, Z = np.meshgrid(ix * dx, -depth + iz * dz)          # (nz+1, nx+1), row-major
    xlocs = X.ravel().astype(np.float64)
    zlocs = Z.ravel().astype(np.float64)
    nnpgx = nx + 1
    ex, ez = np.meshgrid(np.arange(nx), np.arange(nz))
    a1 = (ez * nnpgx + ex).ravel() + 1
    ieng = np.stack([a1, a1 + 1, a1 + 1 + nnpgx, a1 + nnpgx], axis=1).astype(np.int32)

    def layer(i, j):  # distance in element layers from the absorbing outer boundary
        return min(i, nx - 1 - i, j)

    dom = bytearray()
    for e in range(nx * nz):
        i, j = e % nx, e // nx
        d = layer(i, j)
        dom += b'A' if d < npml else (b'E' if d < npml + nbielak else b'I')
    cn = bytearray()
    xd = np.zeros(xlocs.shape)
    zd = np.zeros(xlocs.shape)
    xl, xr = npml * dx, width - npml * dx
    zb = -depth + npml * dz
    for g in range(xlocs.size):
        i, j = g % nnpgx, g // nnpgx
        if j == nz:
            cn += b'FS'
        elif npml > 0 and j == 0:
            cn += b'FB'
        elif npml > 0 and i == 0:
            cn += b'FL'
        elif npml > 0 and i == nx:
            cn += b'FR'
        else:
            d = min(i, nx - i, j)
            if npml > 0 and d < npml:
                cn += b'AA'
            elif npml > 0 and d < npml + nbielak:
                cn += b'EE'
            elif (npml + nbielak) > 0 and d == npml + nbielak:
                cn += b'BI'
            else:
                cn += b'II'
        if npml > 0:
            xd[g] = max(xl - xlocs[g], 0.0, xlocs[g] - xr)
            zd[g] = max(zb - zlocs[g], 0.0)
    return Mesh(nord=nord, nx=nx, nz=nz, ieng=ieng, xlocs=xlocs, zlocs=zlocs,
                cdomain=bytes(dom), cnnpg=bytes(cn), xd=xd, zd=zd,
                xwidth=npml * dx, zwidth=npml * dz)


def model_from_velocity(vp: np.ndarray, vs: np.ndarray, rho: np.ndarray) -> Model:
    lam = rho * (vp * vp - 2.0 * vs * vs)
    mu = rho * vs * vs
    return Model(dens=np.ascontiguousarray(rho, dtype=np.float64),
                 ecoeff=np.ascontiguousarray(np.stack([lam, mu], axis=1), dtype=np.float64),
                 vp=np.asarray(vp, np.float64), vs=np.asarray(vs, np.float64))


def homogeneous_model(mesh: Mesh, vp=6000.0, vs=3460.0, rho=2700.0) -> Model:
    n = mesh.nnpg
    return model_from_velocity(np.full(n, vp), np.full(n, vs), np.full(n, rho))


def layered_model(mesh: Mesh, seed: int = 2, n_anom: int = 20, amp: float = 0.05,
                  anomaly_scale: float = 1.0) -> Model:
    """35 km crust (Vp 6.3 km/s) over mantle (Vp 8.0 km/s), Vs = Vp/sqrt(3),
    rho[g/cc] = 0.32 Vp[km/s] + 0.77, plus n_anom seeded Gaussian +-amp Vp
    anomalies (sigma 5-15 km); Vs and rho are re-derived from the perturbed
    Vp.  anomaly_scale multiplies the anomaly amplitudes."""
    rng = np.random.default_rng(seed)
    x, z = mesh.xlocs, mesh.zlocs
    vp = np.where(z > -35e3, 6300.0, 8000.0)
    xmin, xmax, zmin, zmax = x.min(), x.max(), z.min(), z.max()
    pert = np.zeros_like(vp)
    for _ in range(n_anom):
        cx = rng.uniform(xmin, xmax)
        cz = rng.uniform(zmin, zmax)
        sig = rng.uniform(5e3, 15e3)
        sgn = 1.0 if rng.uniform() < 0.5 else -1.0
        pert += sgn * amp * np.exp(-((x - cx) ** 2 + (z - cz) ** 2) / (2.0 * sig * sig))
    vp = vp * (1.0 + anomaly_scale * pert)
    vs = vp / math.sqrt(3.0)
    rho = (0.32 * vp / 1000.0 + 0.77) * 1000.0
    return model_from_velocity(vp, vs, rho)


def freq_ky_batch(nfreq: int, nky: int, fmin: float, fmax: float, vs_min: float,
                  ky_frac: float = 0.5):
    """(omega, ky, freq, py) arrays for nfreq frequencies uniform in
    [fmin, fmax] and nky wavenumbers uniform in [0, ky_frac * omega / vs_min]
    per frequency; py = ky / omega (estiff25.f:61)."""
    freqs = np.linspace(fmin, fmax, nfreq)
    om, ky = [], []
    for f in freqs:
        w = 2.0 * math.pi * f
        for k in np.linspace(0.0, ky_frac * w / vs_min, nky):
            om.append(w)
            ky.append(k)
    om = np.array(om)
    ky = np.array(ky)
    return om, ky, om / (2.0 * math.pi), ky / om


def seeded_complex(n: int, batch: int | None = None, seed: int = 1) -> np.ndarray:
    rng = np.random.default_rng(seed)
    shape = (n,) if batch is None else (batch, n)
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


# Named workloads (sizes from the benchmark definition).
CONFIGS = {
    "c1": dict(nx=40, nz=20, nord=4, nfreq=1, nky=1),
    "c2": dict(nx=200, nz=100, nord=4, nfreq=8, nky=16),
    "c5p4": dict(nx=400, nz=200, nord=4, nfreq=6, nky=32),
    "c5p5": dict(nx=400, nz=200, nord=5, nfreq=6, nky=32),
    "c5p6": dict(nx=400, nz=200, nord=6, nfreq=6, nky=32),
}
