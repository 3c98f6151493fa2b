"""Seeded structure and GAP-model generators for the benchmark workloads.

Every random choice comes from numpy PCG64 generators seeded from the workload
index (seed = 1000 * workload + k), so a workload is reproducible from its
name alone.  Models are ordinary GAP XML files (quip_amd.gap_xml) written from
the seed, with the sparse points taken as descriptor vectors of perturbed
copies of the training cell.

Workloads (sizes as named in the benchmark specification):
  si64      64-atom diamond Si (a=5.43, 2x2x2), sigma=0.1; SOAP n_max=8 l_max=6
            cutoff 5.0 (d=253) with 500 sparse points + a distance_2b term
  si4096    8x8x8 diamond Si, sigma=0.15, 1% vacancies (4055 atoms);
            same SOAP with 2000 sparse points
  sic32768  16^3 zincblende SiC (a=4.36), sigma=0.1, 2% antisite swaps;
            SOAP n_max=12 l_max=8 two species (d=2701), one coordinate per
            centre species with 2000 sparse points each (4000 total)
  batch512  512 Si structures, N ~ U[8,512], bulk / (111) slab / vacancy mix
"""
import numpy as np

from .atoms import Atoms
from .gap_xml import Descriptor, GAPModel, GPCoordinate, GAP_VERSION

DIAMOND_BASIS = np.array([[0, 0, 0], [0, .5, .5], [.5, 0, .5], [.5, .5, 0],
                          [.25, .25, .25], [.25, .75, .75], [.75, .25, .75], [.75, .75, .25]])


def rng_for(workload, k=0):
    return np.random.Generator(np.random.PCG64(1000 * workload + k))


def diamond(a, reps, species=(14, 14)):
    """Conventional diamond/zincblende supercell: first 4 basis sites species[0], last 4 species[1]."""
    nx, ny, nz = reps
    cells = np.array([[i, j, k] for i in range(nx) for j in range(ny) for k in range(nz)], dtype=float)
    pos = (cells[:, None, :] + DIAMOND_BASIS[None, :, :]).reshape(-1, 3) * a
    Z = np.tile(np.array([species[0]] * 4 + [species[1]] * 4), len(cells))
    return Atoms(Z, pos, np.diag([nx * a, ny * a, nz * a]), (True, True, True))


def perturb(at, sigma, rng):
    at = at.copy()
    at.positions = at.positions + rng.normal(0.0, sigma, at.positions.shape)
    return at


def remove_random(at, count, rng):
    keep = np.sort(rng.choice(len(at), len(at) - count, replace=False))
    return Atoms(at.numbers[keep], at.positions[keep], at.cell, at.pbc)


def antisite_swaps(at, n_pairs, rng):
    at = at.copy()
    a = np.where(at.numbers == at.numbers.max())[0]
    b = np.where(at.numbers == at.numbers.min())[0]
    ia = rng.choice(a, n_pairs, replace=False)
    ib = rng.choice(b, n_pairs, replace=False)
    za, zb = at.numbers[ia].copy(), at.numbers[ib].copy()
    at.numbers[ia], at.numbers[ib] = zb, za
    return at


def si64(seed_k=1):
    return perturb(diamond(5.43, (2, 2, 2)), 0.1, rng_for(0, seed_k))


def si4096():
    rng = rng_for(1, 1)
    at = perturb(diamond(5.43, (8, 8, 8)), 0.15, rng)
    return remove_random(at, 41, rng)


def si4096_batch(n_struct=8):
    """n_struct structures made as si4096 is (8x8x8 diamond Si, sigma 0.15, 41 vacancies), one seed
    each; the first is si4096 itself."""
    out = []
    for k in range(n_struct):
        rng = rng_for(1, 1 + k)
        at = perturb(diamond(5.43, (8, 8, 8)), 0.15, rng)
        out.append(remove_random(at, 41, rng))
    return out


def sic32768():
    rng = rng_for(2, 1)
    at = perturb(diamond(4.36, (16, 16, 16), species=(14, 6)), 0.1, rng)
    return antisite_swaps(at, 328, rng)


def slab111(a, layers, vacuum, reps=(1, 1), species=(14, 14)):
    """Diamond (111) slab of `layers` atomic layers (layers/2 bilayers), in-plane cell
    3w*reps[0] x 2*sqrt(3)w*reps[1] with w = a/sqrt(2) (commensurate with the (111)
    lattice: w along [1-10], sqrt(3)w along [11-2]; 24 atoms per bilayer per unit),
    `vacuum` Angstrom of empty space along z, and z non-periodic.  species as in diamond()."""
    n = np.array([1.0, 1.0, 1.0]) / np.sqrt(3.0)
    u = np.array([1.0, -1.0, 0.0]) / np.sqrt(2.0)
    v = np.cross(n, u)
    w = a / np.sqrt(2.0)
    L1, L2 = 3 * w * reps[0], 2 * np.sqrt(3.0) * w * reps[1]
    d111 = a / np.sqrt(3.0)
    thick = layers * d111 / 2.0
    m = int(np.ceil(1.5 * max(L1, L2, thick) / a)) + 2
    bulk = diamond(a, (m, m, m), species)
    p = (bulk.positions - bulk.positions.mean(axis=0)) @ np.stack([u, v, n]).T
    sel = (p[:, 2] >= -1e-6) & (p[:, 2] < thick - 1e-6)
    p, Z = p[sel], bulk.numbers[sel]
    p[:, 0] = np.mod(p[:, 0], L1)
    p[:, 1] = np.mod(p[:, 1], L2)
    key = np.round(np.mod(p[:, :2] + 1e-6, [L1, L2]) * 1e4).astype(np.int64)
    key = np.column_stack([key, np.round(p[:, 2] * 1e4).astype(np.int64)])
    _, first = np.unique(key, axis=0, return_index=True)
    p, Z = p[np.sort(first)], Z[np.sort(first)]
    assert len(p) == 24 * reps[0] * reps[1] * (layers // 2), "slab block too small"
    p[:, 2] += vacuum / 2.0
    cell = np.diag([L1, L2, thick + vacuum])
    return Atoms(Z, p, cell, (True, True, False))


def batch512(n_struct=512, workload=3):
    """512 Si structures, N ~ U[8, 512]; by seed: periodic bulk, (111) slab with 10 A
    vacuum (z non-periodic), or bulk with 1-3 vacancies; a in [5.3, 5.6], sigma in [0.05, 0.3]."""
    rng = rng_for(workload, 1)
    out = []
    for s in range(n_struct):
        target = int(rng.integers(8, 513))
        kind = int(rng.integers(0, 3))
        a = float(rng.uniform(5.3, 5.6))
        sig = float(rng.uniform(0.05, 0.3))
        out.append(mixed_structure(rng, target, a, sig, kind))
    return out


def mixed_structure(rng, target, a, sigma, kind, species=(14, 14)):
    """One structure of about `target` atoms (fewer after vacancies): kind 0 periodic
    bulk, 1 (111) slab with 10 A vacuum (z non-periodic), 2 bulk with 1-3 vacancies."""
    if kind == 1:
        k = int(np.ceil(target / 72.0))
        r1 = int(np.ceil(np.sqrt(k)))
        at = slab111(a, 6, 10.0, (r1, int(np.ceil(k / r1))), species)
    else:
        reps = max(1, int(round((target / 8.0) ** (1.0 / 3.0))))
        at = diamond(a, (reps, reps, max(1, int(np.ceil(target / (8.0 * reps * reps))))), species)
    if len(at) > target:
        at = remove_random(at, len(at) - target, rng)
    if kind == 2 and len(at) > 8:
        at = remove_random(at, int(rng.integers(1, 4)), rng)
    return perturb(at, sigma, rng)


def sic_requests(n_req, workload=4, n_min=16, n_max=4096):
    """Serving stream: Si/C zincblende structures (a in [4.30, 4.42]), N log-uniform on
    [n_min, n_max], kinds and displacements drawn as for batch512."""
    rng = rng_for(workload, 1)
    out = []
    for _ in range(n_req):
        target = int(round(np.exp(rng.uniform(np.log(n_min), np.log(n_max)))))
        kind = int(rng.integers(0, 3))
        a = float(rng.uniform(4.30, 4.42))
        sig = float(rng.uniform(0.05, 0.3))
        out.append(mixed_structure(rng, target, a, sig, kind, species=(14, 6)))
    return out


def soap_string(cutoff=5.0, ctw=1.0, sigma=0.5, n_max=8, l_max=6, species=(14,), centre=14):
    return ("soap cutoff=%g cutoff_transition_width=%g atom_sigma=%g n_max=%d l_max=%d central_weight=1.0 "
            "n_species=%d species_Z={%s} n_Z=1 Z={%d}" % (cutoff, ctw, sigma, n_max, l_max, len(species),
                                                            " ".join(str(z) for z in species), centre))


def soap_model(descriptor_fn, train, n_sparse, workload, desc_strings, label, zeta=4.0, delta=1.0,
               alpha_std=1e-3, d2b=None):
    """Build a SOAP GAP model: for each descriptor string, candidate sparse points are
    descriptor_fn(Descriptor, Atoms) rows over `train`, picked without replacement."""
    rng = rng_for(workload, 7)
    coords = []
    for k, ds in enumerate(desc_strings):
        d = Descriptor.from_string(ds)
        cand = np.vstack([descriptor_fn(d, at) for at in train])
        pick = np.sort(rng.choice(len(cand), n_sparse, replace=False))
        X = np.ascontiguousarray(cand[pick])
        alpha = rng.normal(0.0, alpha_std, n_sparse)
        coords.append(GPCoordinate("%s%d" % (label, k + 1), d, 2, delta, 0.0, zeta, X, alpha, np.ones(n_sparse)))
    if d2b is not None:
        cutoff, theta, n2, lo, hi = d2b
        d = Descriptor.from_string("distance_2b cutoff=%g cutoff_transition_width=0.5 Z1=14 Z2=14" % cutoff)
        X = np.linspace(lo, hi, n2).reshape(-1, 1)
        alpha = rng.normal(0.0, alpha_std, n2)
        coords.append(GPCoordinate("%s%d" % (label, len(coords) + 1), d, 1, delta, 0.0, 1.0, X, alpha,
                                   np.ones(n2), theta=np.array([theta])))
    return GAPModel(label, GAP_VERSION, np.zeros(119), coords,
                    "generated from seed %d: %s" % (workload, "; ".join(desc_strings)))


def model_si64(descriptor_fn):
    train = [perturb(diamond(5.43, (2, 2, 2)), 0.1, rng_for(0, 100 + k)) for k in range(20)]
    return soap_model(descriptor_fn, train, 500, 0, [soap_string()], "GAP_si64", d2b=(5.0, 1.0, 20, 1.8, 5.0))


def model_si4096(descriptor_fn):
    train = [perturb(diamond(5.43, (2, 2, 2)), 0.15, rng_for(1, 100 + k)) for k in range(32)]
    return soap_model(descriptor_fn, train, 2000, 1, [soap_string()], "GAP_si4096")


def model_sic(descriptor_fn):
    train = [antisite_swaps(perturb(diamond(4.36, (2, 2, 2), species=(14, 6)), 0.1, rng_for(2, 100 + k)), 1,
                            rng_for(2, 200 + k)) for k in range(64)]
    strs = [soap_string(n_max=12, l_max=8, species=(14, 6), centre=14),
            soap_string(n_max=12, l_max=8, species=(14, 6), centre=6)]
    return soap_model(descriptor_fn, train, 2000, 2, strs, "GAP_sic")
