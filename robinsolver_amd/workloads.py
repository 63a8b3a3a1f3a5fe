"""Seeded synthetic inputs of the benchmark configurations, sampled on the GPU.

The reference samples the potential k, the sources f and the boundary data g
through host callbacks at the leaf points (fd_module.f90:136-145, 427-431,
645-651).  These helpers evaluate the same kind of functions with torch on the
device, at the device coordinates that rs_points_device writes (the host's
bits), so that drivers and tests of the large configurations need no host
sampling.  Host (numpy) versions of the potential and sources are kept for
drivers that pass host arrays.

Configurations (domain [0,1]^d, Robin a1 = -i kb realised as opts%kb = -kb):
  CFG1  2D 64^2,   L=4,  kb=10, centred Gaussian bump potential, 1 RHS
  CFG2  2D 512^2,  L=10, kb=40, 16-mode potential + absorption, 64 RHS, seed 2
  CFG3  2D 1024^2, L=12, kb=80, 16-mode potential + absorption, 256 RHS, seed 3
  CFG4  3D 64^3,   L=9,  kb=8,  24-mode potential, 16 RHS (sigma 0.05), seed 4
  MIX_TREES  the eight resident trees of the serving configuration
"""
import numpy as np

CFG2 = dict(d=2, grid=512, L=10, kb=40.0, seed=2, nrhs=64, nmodes=16, sigma=0.02, absorb=0.01)
CFG3 = dict(d=2, grid=1024, L=12, kb=80.0, seed=3, nrhs=256, nmodes=16, sigma=0.02, absorb=0.01)
CFG4 = dict(d=3, grid=64, L=9, kb=8.0, seed=4, nrhs=16, nmodes=24, sigma=0.05, absorb=0.0)

# (grid, L, kb): L gives leaves of about 16 x 16 points (12, 14, 16 or 20 per side)
MIX_TREES = [(128, 6, 10.0), (256, 8, 10.0), (384, 10, 20.0), (512, 10, 20.0), (640, 10, 40.0), (768, 12, 40.0),
             (896, 12, 80.0), (1024, 12, 80.0)]


def _modes(seed, nmodes, d):
    rng = np.random.default_rng(seed)
    a = rng.uniform(-1, 1, nmodes)
    p = rng.integers(1, 9, size=(nmodes, d))
    ph = rng.uniform(0, 2 * np.pi, nmodes)
    return a, p, ph


def potential(pts_int, kb, seed, nmodes=16, d=2, absorb=0.01):
    """kb^2 (1 + 0.3 q) + i absorb kb^2 at host points [..., 3],
    q = sum_m a_m cos(2 pi p_m.x + phi_m) / sum|a_m|, a_m ~ U[-1,1],
    p_m in {1..8}^d, phi_m ~ U[0, 2 pi)."""
    a, p, ph = _modes(seed, nmodes, d)
    q = np.zeros(pts_int.shape[:-1])
    for j in range(nmodes):
        q += a[j] * np.cos(2 * np.pi * (pts_int[..., :d] @ p[j]) + ph[j])
    q /= np.abs(a).sum()
    return kb ** 2 * (1 + 0.3 * q) + 1j * absorb * kb ** 2


def potential_dev(torch, pts_int_dev, kb, seed, nmodes=16, d=2, absorb=0.01):
    """potential() on the GPU at device coordinates (the same seeded modes;
    complex128 [n_leaves][n_int])."""
    a, p, ph = _modes(seed, nmodes, d)
    x = pts_int_dev[..., :d]
    q = torch.zeros(x.shape[:-1], dtype=torch.float64, device=x.device)
    for j in range(nmodes):
        pj = torch.tensor(p[j], dtype=torch.float64, device=x.device)
        q += float(a[j]) * torch.cos(2 * np.pi * (x @ pj) + float(ph[j]))
    q /= float(np.abs(a).sum())
    return (kb ** 2 * (1 + 0.3 * q)).to(torch.complex128) + 1j * absorb * kb ** 2


def device_points(torch, tree):
    """Every point's coordinates [n_leaves][m][3] and face id [n_leaves][m] on the GPU."""
    xyz = torch.empty((tree.n_leaves, tree.m, 3), dtype=torch.float64, device="cuda:0")
    face = torch.empty((tree.n_leaves, tree.m), dtype=torch.int32, device="cuda:0")
    tree.points_device(xyz.data_ptr(), face.data_ptr())
    torch.cuda.synchronize()
    return xyz, face


def source_centres(nrhs, seed, d=2):
    """Centres of the Gaussian sources, U[0.1, 0.9]^d."""
    rng = np.random.default_rng(seed + 1000)
    return rng.uniform(0.1, 0.9, size=(nrhs, d))


def gaussian_sources(torch, pts_int_dev, nrhs, seed, sigma=0.02, d=2, first=0):
    """f_r = exp(-|x - c_r|^2 / (2 sigma^2)) for r in [first, first + nrhs) of the
    seeded centre sequence; complex128 [nrhs][n_leaves][n_int] on the GPU."""
    c_all = source_centres(first + nrhs, seed, d)[first:]
    c = torch.tensor(c_all, device=pts_int_dev.device, dtype=torch.float64)
    x = pts_int_dev[..., :d]
    f = torch.empty((nrhs,) + tuple(x.shape[:2]), dtype=torch.complex128, device=x.device)
    for r in range(nrhs):
        r2 = ((x - c[r]) ** 2).sum(-1)
        f[r] = torch.exp(-r2 / (2 * sigma * sigma)).to(torch.complex128)
    return f


def plane_wave_g_dev(torch, tree, xyz, face, kb, thetas):
    """Impedance data of u = exp(i kb d.x), d = (cos t, sin t), on the physical
    faces of a 2D tree: g = a1 u + du/dn with a1 = -i kb, so
    g = i kb (d.n - 1) u; evaluated on the GPU from device coordinates and face
    ids.  Returns [len(thetas)][n_leaves][n_bnd] (zero away from the boundary)."""
    nb = tree.nb
    x = xyz[:, :nb, :2]
    fc = face[:, :nb]
    normal = torch.zeros_like(x)
    phys = torch.zeros(x.shape[:2], dtype=torch.bool, device=x.device)
    for fid, ax, sgn, edge in ((1, 0, -1.0, 0.0), (2, 0, 1.0, 1.0), (3, 1, -1.0, 0.0), (4, 1, 1.0, 1.0)):
        on = (fc == fid) & ((x[..., ax] - edge).abs() < 1e-12)
        phys |= on
        normal[..., ax] = torch.where(fc == fid, torch.full_like(normal[..., ax], sgn), normal[..., ax])
    out = torch.zeros((len(thetas),) + tuple(x.shape[:2]), dtype=torch.complex128, device=x.device)
    for i, th in enumerate(thetas):
        dv = torch.tensor([np.cos(th), np.sin(th)], dtype=torch.float64, device=x.device)
        u = torch.exp(1j * kb * (x @ dv).to(torch.complex128))
        val = 1j * kb * ((normal @ dv) - 1.0).to(torch.complex128) * u
        out[i] = torch.where(phys, val, torch.zeros_like(val))
    return out


# Comparisons of device results are made on the host: torch's complex norm
# reductions on this ROCm build have returned wrong values now and then (a rel.
# l2 of 0.22 for two arrays 1e-11 apart), so a decision never rests on them.
def _host(x):
    return x.detach().cpu().numpy() if hasattr(x, "detach") else np.asarray(x)


def rel_l2_cols(u, ref, chunk=8):
    """rel. l2 per right-hand side, ||u_r - ref_r|| / ||ref_r|| over each
    column's leaf vectors (get_solution order); host arithmetic, `chunk` columns
    at a time."""
    n = u.shape[0]
    out = np.zeros(n)
    for c0 in range(0, n, chunk):
        a = _host(u[c0:c0 + chunk]).reshape(min(chunk, n - c0), -1)
        b = _host(ref[c0:c0 + chunk]).reshape(a.shape[0], -1)
        out[c0:c0 + a.shape[0]] = np.linalg.norm(a - b, axis=1) / np.linalg.norm(b, axis=1)
    return out


def same_bits(a, b):
    """Bitwise equality of two arrays (host comparison)."""
    x, y = _host(a), _host(b)
    return x.shape == y.shape and x.dtype == y.dtype and np.array_equal(x.view(np.uint8), y.view(np.uint8))

