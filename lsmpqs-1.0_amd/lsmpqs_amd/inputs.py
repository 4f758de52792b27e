"""Seeded synthetic porous media in the reference's input layout.

The reference takes a segmented image, turns it into a signed-distance mask
(positive in solid) and starts the non-wetting phase as a large sphere centred
outside the x-low face (reference: src/applications/createMaskUBC.c:90-116,
src/applications/createInitFromMask.c:80-104).  Here the image is replaced by
a Boolean sphere pack (centres uniform in the fill box, overlap allowed,
spheres added until the voxel porosity reaches the target) whose mask is the
exact union signed distance max_l (r_l - |x - c_l|), clipped at -clip in the
pore space.  Arrays are (nz, ny, nx) views of the column-major ghost box.
"""
import numpy as np


def ghost_coords(grid):
    """Physical x, y, z of every ghost-box voxel centre, as in
    createIntersectionOfSpheres (reference: src/geometry/qss_initialization3d.c:136-141)."""
    out = []
    for a in range(3):
        n = grid.grid_dims_ghostbox[a]
        out.append(grid.x_lo_ghostbox[a] + grid.dx[a] * np.arange(n, dtype=np.float64))
    return out


def sphere_pack_mask(grid, radius, porosity, seed, clip=6.0, max_spheres=10_000_000,
                     device=None):
    """Signed distance to a Boolean sphere pack, > 0 in solid, float64.

    radius: a number, or a (rmin, rmax) pair for radii uniform in [rmin, rmax].
    device: None builds the mask with numpy; a torch device ("cuda") builds
    it there from the same seeded spheres (for packs of 1e9 voxels) and
    returns it to the host.
    Returns (mask, centres (m,3), radii (m,), achieved fill-box porosity)."""
    if device is not None:
        return _sphere_pack_mask_torch(grid, radius, porosity, seed, clip, max_spheres, device)
    rng = np.random.default_rng(seed)
    xs, ys, zs = ghost_coords(grid)
    dx = float(grid.dx[0])
    N = [grid.grid_dims_ghostbox[a] for a in range(3)]
    mask = np.full((N[2], N[1], N[0]), -clip, dtype=np.float64)
    lo = [float(grid.x_lo[a]) for a in range(3)]
    hi = [float(grid.x_hi[a]) for a in range(3)]
    fb = (slice(grid.klo_fb, grid.khi_fb + 1), slice(grid.jlo_fb, grid.jhi_fb + 1),
          slice(grid.ilo_fb, grid.ihi_fb + 1))
    if np.isscalar(radius):
        rmin = rmax = float(radius)
    else:
        rmin, rmax = float(radius[0]), float(radius[1])
    vol_box = np.prod([hi[a] - lo[a] for a in range(3)])
    er3 = (rmax ** 4 - rmin ** 4) / (4 * (rmax - rmin)) if rmax > rmin else rmin ** 3
    expected = int(np.ceil(-np.log(porosity) * vol_box / (4.0 / 3.0 * np.pi * er3)))
    centres, radii = [], []

    def add(count):
        for _ in range(count):
            c = np.array([rng.uniform(lo[a], hi[a]) for a in range(3)])
            r = rmin if rmax == rmin else rng.uniform(rmin, rmax)
            centres.append(c)
            radii.append(r)
            reach = r + clip
            sl = []
            for a, axis in enumerate((xs, ys, zs)):
                i0 = max(0, int(np.floor((c[a] - reach - axis[0]) / dx)))
                i1 = min(N[a], int(np.ceil((c[a] + reach - axis[0]) / dx)) + 1)
                sl.append((i0, i1))
            if any(i1 <= i0 for i0, i1 in sl):
                continue
            X = xs[sl[0][0]:sl[0][1]] - c[0]
            Y = ys[sl[1][0]:sl[1][1]] - c[1]
            Z = zs[sl[2][0]:sl[2][1]] - c[2]
            d = r - np.sqrt(Z[:, None, None] ** 2 + Y[None, :, None] ** 2 + X[None, None, :] ** 2)
            view = mask[sl[2][0]:sl[2][1], sl[1][0]:sl[1][1], sl[0][0]:sl[0][1]]
            np.maximum(view, d, out=view)

    add(int(0.9 * expected))
    step = max(1, expected // 50)
    while True:
        phi_now = float(np.mean(mask[fb] < 0))
        if phi_now <= porosity or len(radii) >= max_spheres:
            break
        add(step)
    return mask, np.array(centres), np.array(radii), phi_now


def _sphere_pack_mask_torch(grid, radius, porosity, seed, clip, max_spheres, device):
    """sphere_pack_mask with the mask held and updated as a torch tensor on
    `device`; the spheres come from the same seeded generator in the same
    order."""
    import torch
    rng = np.random.default_rng(seed)
    coords = ghost_coords(grid)
    xs, ys, zs = (torch.from_numpy(c).to(device) for c in coords)
    axes_np = coords
    dx = float(grid.dx[0])
    N = [grid.grid_dims_ghostbox[a] for a in range(3)]
    mask = torch.full((N[2], N[1], N[0]), -clip, dtype=torch.float64, device=device)
    lo = [float(grid.x_lo[a]) for a in range(3)]
    hi = [float(grid.x_hi[a]) for a in range(3)]
    fb = (slice(grid.klo_fb, grid.khi_fb + 1), slice(grid.jlo_fb, grid.jhi_fb + 1),
          slice(grid.ilo_fb, grid.ihi_fb + 1))
    rmin, rmax = (float(radius), float(radius)) if np.isscalar(radius) else \
        (float(radius[0]), float(radius[1]))
    vol_box = np.prod([hi[a] - lo[a] for a in range(3)])
    er3 = (rmax ** 4 - rmin ** 4) / (4 * (rmax - rmin)) if rmax > rmin else rmin ** 3
    expected = int(np.ceil(-np.log(porosity) * vol_box / (4.0 / 3.0 * np.pi * er3)))
    centres, radii = [], []

    def add(count):
        for _ in range(count):
            c = np.array([rng.uniform(lo[a], hi[a]) for a in range(3)])
            r = rmin if rmax == rmin else rng.uniform(rmin, rmax)
            centres.append(c)
            radii.append(r)
            reach = r + clip
            sl = []
            for a in range(3):
                i0 = max(0, int(np.floor((c[a] - reach - axes_np[a][0]) / dx)))
                i1 = min(N[a], int(np.ceil((c[a] + reach - axes_np[a][0]) / dx)) + 1)
                sl.append((i0, i1))
            if any(i1 <= i0 for i0, i1 in sl):
                continue
            X = xs[sl[0][0]:sl[0][1]] - c[0]
            Y = ys[sl[1][0]:sl[1][1]] - c[1]
            Z = zs[sl[2][0]:sl[2][1]] - c[2]
            d = r - torch.sqrt(Z[:, None, None] ** 2 + Y[None, :, None] ** 2
                               + X[None, None, :] ** 2)
            view = mask[sl[2][0]:sl[2][1], sl[1][0]:sl[1][1], sl[0][0]:sl[0][1]]
            torch.maximum(view, d, out=view)

    add(int(0.9 * expected))
    step = max(1, expected // 50)
    while True:
        phi_now = float((mask[fb] < 0).sum().item()) / float(mask[fb].numel())
        if phi_now <= porosity or len(radii) >= max_spheres:
            break
        add(step)
    out = mask.cpu().numpy()
    del mask
    return out, np.array(centres), np.array(radii), phi_now


def initial_phi(grid, mask):
    """Non-wetting phase as a sphere of radius 0.7 L centred 0.49 L before the
    x-low face, masked by the solid (reference:
    src/applications/createInitFromMask.c:80-86,104)."""
    xs, ys, zs = ghost_coords(grid)
    L = float(grid.x_hi[0] - grid.x_lo[0])
    r = 0.7 * L
    cx = float(grid.x_lo[0]) - r * 0.7
    cy = 0.5 * float(grid.x_lo[1] + grid.x_hi[1])
    cz = 0.5 * float(grid.x_lo[2] + grid.x_hi[2])
    d = np.sqrt((zs[:, None, None] - cz) ** 2 + (ys[None, :, None] - cy) ** 2
                + (xs[None, None, :] - cx) ** 2) - r
    return np.maximum(d, mask)


def pack_case(dims, radius, porosity, seed, dtype=np.float64, clip=6.0):
    """(grid, mask, phi0, info) for one seeded pack in run precision dtype."""
    from .abi import make_grid
    g = make_grid(dims, dtype=dtype)
    mask, c, r, achieved = sphere_pack_mask(g, radius, porosity, seed, clip=clip)
    phi0 = initial_phi(g, mask)
    info = {"spheres": int(len(r)), "porosity": achieved, "seed": int(seed)}
    return g, mask.astype(dtype), phi0.astype(dtype), info


def gaussian_random_field(shape, sigma, corr_len, seed):
    """Stationary Gaussian field of standard deviation sigma whose correlation
    falls as exp(-r^2 / (2 corr_len^2)) (r in voxels): seeded white noise
    filtered with a periodic Gaussian of width corr_len / sqrt(2), rescaled to
    unit sample deviation, times sigma."""
    from scipy.ndimage import gaussian_filter
    w = np.random.default_rng(seed).standard_normal(shape)
    f = gaussian_filter(w, corr_len / np.sqrt(2.0), mode="wrap")
    f -= f.mean()
    return sigma * f / f.std()


def contact_angle_case(n=128, radius=8.0, porosity=0.4, seed=1005, mean_deg=40.0,
                       sigma_deg=5.0, corr_len=10.0, dtype=np.float64):
    """(grid, mask, phi0, theta): an n^3 pack with a per-voxel contact angle
    (radians) mean_deg + a Gaussian random field (sigma_deg, corr_len voxels),
    both seeded by `seed`."""
    g, mask, phi0, _ = pack_case((n, n, n), radius, porosity, seed, dtype)
    deg = mean_deg + gaussian_random_field(mask.shape, sigma_deg, corr_len, seed)
    return g, mask, phi0, (np.deg2rad(deg)).astype(dtype)


def _stamp(mask, axes, c, r, clip, dx):
    """mask = max(mask, r - |x - c|) over the voxels within r + clip of c."""
    reach = r + clip
    sl = []
    for a, axis in enumerate(axes):
        i0 = max(0, int(np.floor((c[a] - reach - axis[0]) / dx)))
        i1 = min(len(axis), int(np.ceil((c[a] + reach - axis[0]) / dx)) + 1)
        if i1 <= i0:
            return
        sl.append((i0, i1))
    X = axes[0][sl[0][0]:sl[0][1]] - c[0]
    Y = axes[1][sl[1][0]:sl[1][1]] - c[1]
    Z = axes[2][sl[2][0]:sl[2][1]] - c[2]
    d = r - np.sqrt(Z[:, None, None] ** 2 + Y[None, :, None] ** 2 + X[None, None, :] ** 2)
    view = mask[sl[2][0]:sl[2][1], sl[1][0]:sl[1][1], sl[0][0]:sl[0][1]]
    np.maximum(view, d, out=view)


def boolean_pack(extent, radius, porosity, rng):
    """Centres (m, 3) uniform in [0, extent)^3 and radii of a Boolean sphere
    model whose expected porosity is `porosity`: m = -ln(porosity) V / E[4/3 pi r^3].
    radius: a number or an (rmin, rmax) pair (radii uniform in it)."""
    rmin, rmax = (float(radius), float(radius)) if np.isscalar(radius) else map(float, radius)
    er3 = (rmax ** 4 - rmin ** 4) / (4 * (rmax - rmin)) if rmax > rmin else rmin ** 3
    m = int(round(-np.log(porosity) * float(extent) ** 3 / (4.0 / 3.0 * np.pi * er3)))
    centres = rng.uniform(0.0, float(extent), size=(m, 3))
    radii = rng.uniform(rmin, rmax, size=m) if rmax > rmin else np.full(m, rmin)
    return centres, radii


def subvolume_mask(grid, origin, centres, radii, clip=6.0):
    """Mask of the subvolume of a pack whose fill box starts at parent
    coordinate `origin` (x, y, z): the union signed distance of the spheres
    that reach its ghost box, clipped at -clip, float64."""
    xs, ys, zs = ghost_coords(grid)
    dx = float(grid.dx[0])
    shift = np.array([float(grid.x_lo[a]) - origin[a] for a in range(3)])
    c = centres + shift                      # sphere centres in the grid's coordinates
    lo = np.array([xs[0], ys[0], zs[0]])
    hi = np.array([xs[-1], ys[-1], zs[-1]])
    reach = radii + clip
    keep = np.all((c + reach[:, None] >= lo) & (c - reach[:, None] <= hi), axis=1)
    N = [grid.grid_dims_ghostbox[a] for a in range(3)]
    mask = np.full((N[2], N[1], N[0]), -clip, dtype=np.float64)
    for cc, r in zip(c[keep], radii[keep]):
        _stamp(mask, (xs, ys, zs), cc, r, clip, dx)
    return mask


def subvolume_batch(nbox=64, parent=768, lo=48, hi=192, radius=(6.0, 12.0), porosity=0.36,
                    seed=1004, dtype=np.float32, b=0.04, kappa_start=(0.03, 0.08),
                    kappa_end=(0.2, 0.4), clip=6.0):
    """Subvolumes cut from one seeded parent^3 Boolean pack: per box its grid
    (dims uniform in {lo..hi} per axis), mask, initial phi (the non-wetting
    sphere at its own x-low face) and curvature schedule (kappa = a0/b from
    U(kappa_start) to U(kappa_end)).  Returns a list of dicts."""
    from .abi import make_grid
    rng = np.random.default_rng(seed)
    centres, radii = boolean_pack(parent, radius, porosity, rng)
    out = []
    for _ in range(nbox):
        dims = [int(v) for v in rng.integers(lo, hi + 1, size=3)]
        origin = [int(rng.integers(0, parent - d + 1)) for d in dims]
        g = make_grid(dims, dtype=dtype)
        mask = subvolume_mask(g, origin, centres, radii, clip)
        phi0 = initial_phi(g, mask)
        ks, ke = rng.uniform(*kappa_start), rng.uniform(*kappa_end)
        out.append({"grid": g, "dims": dims, "origin": origin, "mask": mask.astype(dtype),
                    "phi0": phi0.astype(dtype), "kappa": (ks, ke), "b": b})
    return out
