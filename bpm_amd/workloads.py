"""Synthetic inputs of the benchmark workloads (seeded numpy).

* sphere volumes: dn(r) = sum over spheres of dn_s [|r - c_s| < r_s] on voxel
  positions (i dx, j dy, k dz), overlaps summed;
* the plain-BPM workloads: a 256^2 x 256 volume with one centred r=2 um
  dn=0.05 sphere plus 20 spheres r~U(.5,1.5) um, dn~U(.01,.03), and a
  1024^2 x 1024 volume of 400 spheres r~U(1,4) um, dn~U(.01,.05);
* a paraxial Gaussian beam (waist w0, focus at depth zf) and tilted plane
  waves, both in the exp(-i k z) convention of the propagator.
"""
import numpy as np

__all__ = ["random_spheres", "small_workload_spheres", "spheres_dn", "gaussian_beam",
           "plane_wave", "plane_waves", "plain_small_workload", "plain_large_spheres",
           "plain_large_workload"]


def random_spheres(rng, size, units, n, r_range, dn_range):
    """n spheres with centres uniform in the volume; returns (centres (n,3) xyz um, radii, dns)."""
    Nx, Ny, Nz = size
    dx, dy, dz = units
    ext = np.array([Nx * dx, Ny * dy, Nz * dz])
    c = rng.uniform(0, 1, (n, 3)) * ext
    r = rng.uniform(r_range[0], r_range[1], n)
    d = rng.uniform(dn_range[0], dn_range[1], n)
    return c, r, d


def small_workload_spheres(rng, size, units):
    """Spheres of the 256^2 x 256 workload's recipe on any grid: one centred
    sphere r = 2 um, dn = .05, then 20 spheres with centres uniform in the
    volume, r ~ U(.5, 1.5) um, dn ~ U(.01, .03) (also the recipe of the
    served and differentiable workloads); returns (centres, radii, dns)."""
    Nx, Ny, Nz = size
    ext = np.array([Nx * units[0], Ny * units[1], Nz * units[2]])
    c, r, d = random_spheres(rng, size, units, 20, (.5, 1.5), (.01, .03))
    return (np.vstack([ext / 2, c]), np.concatenate([[2.], r]), np.concatenate([[.05], d]))


def plane_waves(Nx, Ny, dx, dy, lam, n0, thetas):
    """(B, Ny, Nx) complex64 tilted plane waves exp(-i k0 n0 (x sin tx + y sin ty))
    for thetas (B, 2) = (tx, ty) in radians."""
    th = np.asarray(thetas, np.float64).reshape(-1, 2)
    k = 2. * np.pi * n0 / lam
    x = dx * np.arange(Nx)
    y = dy * np.arange(Ny)
    ph = k * (np.sin(th[:, 0])[:, None, None] * x[None, None, :] +
              np.sin(th[:, 1])[:, None, None] * y[None, :, None])
    return np.exp(-1.j * ph).astype(np.complex64)


def spheres_dn(size, units, centres, radii, dns, dtype=np.float32):
    """(Nz,Ny,Nx) volume, sum of dn_s inside each sphere (strict |r-c|<r)."""
    Nx, Ny, Nz = size
    dx, dy, dz = units
    out = np.zeros((Nz, Ny, Nx), dtype)
    for (cx, cy, cz), r, d in zip(centres, radii, dns):
        lo = [max(0, int(np.floor((c - r) / u))) for c, u in zip((cx, cy, cz), (dx, dy, dz))]
        hi = [min(N, int(np.ceil((c + r) / u)) + 1) for c, u, N in
              zip((cx, cy, cz), (dx, dy, dz), (Nx, Ny, Nz))]
        if any(h <= l for l, h in zip(lo, hi)):
            continue
        x = np.arange(lo[0], hi[0]) * dx - cx
        y = np.arange(lo[1], hi[1]) * dy - cy
        z = np.arange(lo[2], hi[2]) * dz - cz
        m = (z[:, None, None] ** 2 + y[None, :, None] ** 2 + x[None, None, :] ** 2) < r * r
        out[lo[2]:hi[2], lo[1]:hi[1], lo[0]:hi[0]] += (d * m).astype(dtype)
    return out


def gaussian_beam(Nx, Ny, dx, dy, lam, n0, w0, z):
    """Paraxial Gaussian beam (waist w0) at axial distance z from its focus,
    centred in the grid; exp(-i(k z + k r^2/2R - gouy)), k = 2 pi n0/lam."""
    x = dx * (np.arange(Nx) - Nx / 2)
    y = dy * (np.arange(Ny) - Ny / 2)
    Y, X = np.meshgrid(y, x, indexing="ij")
    r2 = X ** 2 + Y ** 2
    k = 2. * np.pi * n0 / lam
    zr = np.pi * w0 ** 2 * n0 / lam
    w = w0 * np.sqrt(1 + (z / zr) ** 2)
    inv_R = z / (z ** 2 + zr ** 2)
    gouy = np.arctan(z / zr)
    u = w0 / w * np.exp(-r2 / w ** 2) * np.exp(-1.j * (k * z + k * r2 * inv_R / 2. - gouy))
    return u.astype(np.complex64)


def plane_wave(Nx, Ny, dx, dy, lam, n0, theta_x=0., theta_y=0.):
    """exp(-i k0 n0 (x sin theta_x + y sin theta_y)) on the (Ny,Nx) grid."""
    k = 2. * np.pi * n0 / lam
    x = dx * np.arange(Nx)
    y = dy * np.arange(Ny)
    Y, X = np.meshgrid(y, x, indexing="ij")
    return np.exp(-1.j * k * (X * np.sin(theta_x) + Y * np.sin(theta_y))).astype(np.complex64)


def plain_small_workload(seed=0, size=(256, 256, 256), units=(.1, .1, .1)):
    """dn and u0 of the 256^2 x 256 plain-BPM workload (plane wave at normal incidence)."""
    rng = np.random.default_rng(seed)
    Nx, Ny, Nz = size
    dn = spheres_dn(size, units, *small_workload_spheres(rng, size, units))
    u0 = np.ones((Ny, Nx), np.complex64)
    return u0, dn


def plain_large_spheres(seed=0, size=(1024, 1024, 1024), units=(.1, .1, .1)):
    """Sphere list (centres, radii, dns) of the 1024^2 x 1024 plain-BPM workload."""
    rng = np.random.default_rng(seed)
    return random_spheres(rng, size, units, 400, (1., 4.), (.01, .05))


def plain_large_workload(seed=0, size=(1024, 1024, 1024), units=(.1, .1, .1), lam=.5, n0=1.33,
                         planes=None):
    """dn and u0 of the 1024^2 x 1024 plain-BPM workload: 400 spheres and a
    Gaussian beam (w0 = 3 um) focused 20 um into the volume.  ``planes``
    limits dn to its first planes (same values as the full volume's)."""
    c, r, d = plain_large_spheres(seed, size, units)
    nz = size[2] if planes is None else planes
    dn = spheres_dn((size[0], size[1], nz), units, c, r, d)
    u0 = gaussian_beam(size[0], size[1], units[0], units[1], lam, n0, 3., -20.)
    return u0, dn
