"""Writes the golden fixtures of this directory from the CPU oracle.

The reference ships no golden vectors, and its Vulkan/GLSL path cannot run
here, so each fixture holds a small seeded input and what the oracle (the
CPU restatement of the reference's shaders, itself pinned by the reference's
KATs and closed forms in test_oracle_kats.py) computes for it:

  substep_elastic_apic.npz   10 substeps, one elastic material, APIC, the
                             reference's SVD convention (FP32 and FP64 runs)
  substep_dp_pic.npz         10 substeps, elastic + Drucker-Prager, PIC,
                             proper polar SVD (FP32 and FP64 runs)
  bins_edge.npz              cell keys, stable order, counts and END offsets
                             with particles outside the domain and a crowded cell
  mesh_box_sdf.npz           the mesh SDF of a box, both closest-point variants
  noise.npz                  value noise, fBm and curl at sample points
  emit.npz                   three emission calls of three emitters
  hash_pbf.npz               the PBF / cloth spatial hash (counts, END offsets,
                             sorted indices) with unused slots, particles
                             outside the domain and a crowded cell
  forces_grid.npz            the grid update with gravity, turbulence and
                             noise forces on a P2G grid (occupied nodes only),
                             and the same update with gravity alone

Run from the repository root:  python tests/golden/make_golden.py
Every array is saved with numpy (no pickled objects)."""
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

import oracle  # noqa: E402
from nebula_amd import mesh, mpm  # noqa: E402
from nebula_amd.scenes import _fields, _material, _seed_cells  # noqa: E402


def _block(res, lo, hi, ppc, seed):
    g = [np.arange(lo[a], hi[a]) for a in range(3)]
    cx, cy, cz = np.meshgrid(*g, indexing="ij")
    cells = np.stack([cx.ravel(), cy.ravel(), cz.ravel()], axis=1)
    return _seed_cells(cells, 1.0 / res, ppc, np.random.default_rng(seed))


def _run(inp, precision):
    """The oracle's substeps of a stored case (inputs as saved)."""
    res = int(inp["res"])
    h = 1.0 / res
    sc = oracle.make_scene(float(inp["dt"]), h, (res, res, res), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0),
                           float(inp["flip_ratio"]))
    forces = mpm.forces_bytes([mpm.GravityForce(tuple(inp["gravity_dir"]), float(inp["gravity"]))])
    sim = oracle.Sim(sc, bytes(inp["materials"]), b"", forces, None, precision)
    sim.set_corrected_svd(bool(inp["corrected_svd"]))
    f = inp["fields"]
    if precision == 64:
        sim.set_state_f64(f.astype(np.float64), f[1, :, 3].view(np.uint32))
    else:
        sim.set_state(f)
    sim.substeps(int(inp["substeps"]))
    return sim.get_state_f64() if precision == 64 else sim.get_state()


def substep_cases():
    cases = {}
    # the elastic block of the 64^3 benchmark at 16^3: E = 1e4, nu = 0.2, rho = 1
    res = 16
    pos = _block(res, (6, 8, 6), (10, 12, 10), 8, 1)
    rng = np.random.default_rng(1)
    vel = np.zeros_like(pos)
    vel[:, 1] = -1.0
    vel += rng.normal(0.0, 0.01, vel.shape)
    f = _fields(pos, vel, (1.0 / res) ** 3 / 8)
    mats = [_material(1e4, 0.2, 1.0, mpm.MaterialModel.ELASTIC)]
    cases["substep_elastic_apic"] = dict(
        fields=f, materials=np.frombuffer(mpm.materials_bytes(mats), np.uint8), res=res, dt=1e-4, flip_ratio=-1.0,
        corrected_svd=0, gravity_dir=np.array([0.0, 1.0, 0.0]), gravity=-9.8, substeps=10)
    # an elastic half and a Drucker-Prager half with a deformed start
    res = 24
    pos = _block(res, (8, 9, 8), (16, 13, 16), 4, 2)
    n = pos.shape[0]
    rng = np.random.default_rng(2)
    f = _fields(pos, 0.3 * rng.standard_normal((n, 3)), (1.0 / res) ** 3 / 4,
                (pos[:, 0] >= 0.5).astype(np.uint32))
    for c in range(3):
        f[2 + c, :, :3] = np.eye(3, dtype=np.float32)[c] + 0.05 * rng.standard_normal((n, 3))
    f[2:8, :, 3] = 20.0 * rng.standard_normal((6, n))
    mats = [_material(3e4, 0.3, 1e3, mpm.MaterialModel.ELASTIC),
            _material(3e4, 0.3, 1e3, mpm.MaterialModel.DRUCKER_PRAGER, M=math.tan(math.radians(30)), qc=10.0)]
    cases["substep_dp_pic"] = dict(
        fields=f, materials=np.frombuffer(mpm.materials_bytes(mats), np.uint8), res=res, dt=1e-4, flip_ratio=0.0,
        corrected_svd=1, gravity_dir=np.array([0.0, 1.0, 0.0]), gravity=-9.8, substeps=10)
    for name, inp in cases.items():
        inp = {k: np.asarray(v) for k, v in inp.items()}
        inp["state32"] = _run(inp, 32)
        inp["state64"] = _run(inp, 64)
        yield name, inp


def bins_case():
    res = 16
    pos = _block(res, (4, 4, 4), (12, 12, 12), 2, 3)
    n = pos.shape[0]
    rng = np.random.default_rng(3)
    pos[: n // 8] = rng.uniform(-0.5, 1.5, (n // 8, 3))       # outside the domain: clamped keys
    pos[n // 8: n // 8 + 40] = (0.51, 0.52, 0.53)             # one crowded cell
    pos[-1] = (1.0, 1.0, 1.0)                                  # on the upper boundary
    pos = pos[rng.permutation(n)]                              # storage order is not cell order
    f = _fields(pos.astype(np.float32), np.zeros_like(pos), (1.0 / res) ** 3 / 2)
    sc = oracle.make_scene(1e-4, 1.0 / res, (res, res, res))
    sim = oracle.Sim(sc, mpm.materials_bytes([_material(1e4, 0.2, 1.0, mpm.MaterialModel.ELASTIC)]))
    sim.set_state(f)
    sim.bin()
    keys, srt, cnt, end = sim.get_bins()
    return dict(fields=f, res=np.asarray(res), keys=keys, sorted=srt, cell_count=cnt, cell_end=end)


def mesh_case():
    s = np.sqrt(1.0 / 3.0) * np.sin(0.3)
    tris = mesh.transform_triangles(mesh.box_mesh((-0.2, -0.1, -0.15), (0.2, 0.1, 0.15)), (0.5, 0.45, 0.5),
                                    (np.cos(0.3), s, s, s))
    res, h = 16, 1.0 / 16
    return dict(tris=tris, res=np.asarray(res), sdf_reference=oracle.mesh_sdf(tris, (res,) * 3, h),
                sdf_exact=oracle.mesh_sdf(tris, (res,) * 3, h, exact=True))


def noise_case():
    rng = np.random.default_rng(4)
    pts = rng.uniform(-3.0, 3.0, (64, 3)).astype(np.float32)
    val = np.array([oracle.noise_value(p) for p in pts], np.float32)
    fbm = np.array([oracle.noise_fbm(p, 3) for p in pts], np.float32)
    curl = np.array([oracle.noise_curl(p, 2) for p in pts], np.float32)
    return dict(points=pts, value=val, fbm3=fbm, curl2=curl)


def emit_case():
    a = mpm.AABBEmitter.FromAABB((0.2, 0.5, 0.2), (0.4, 0.7, 0.4), (0.0, -1.0, 0.0), 60)
    a.step_count, a.center_vel = 0, (0.5, 0.0, 0.0)
    sph = mpm.SphereEmitter.FromSphere((0.7, 0.6, 0.6), 0.08, (0.0, 0.0, 2.0), 40)
    sph.emitAlongNormal, sph.step_count, sph.particleType = True, 2, 1
    el = mpm.EllipseEmitter()
    el.center, el.semiA, el.semiB, el.particles_per_step = (0.5, 0.3, 0.5), 0.1, 0.05, 30
    el.velocityRandomness, el.step_count = 0.3, -1
    recs = b"".join(bytes(x.pack()) for x in (a, sph, el))
    Vp = float(np.float32(1.0 / 8) ** 3)
    pos, vel = oracle.emit(recs, 3, 3, 0.01, 1000, 0, Vp)
    return dict(records=np.frombuffer(recs, np.uint8), n_emitters=np.asarray(3), calls=np.asarray(3),
                dt=np.asarray(0.01), capacity=np.asarray(1000), Vp=np.asarray(Vp), pos=pos, vel=vel)


def forces_grid_case():
    """A P2G grid of a moving block, then the grid update with gravity,
    turbulence (curl of fBm) and noise (fBm along a direction), and with
    gravity alone (so a test can see that the other two forces acted).  Only
    nodes with mass >= 1e-10 are stored: their Morton ids, the input
    (momentum, mass) and the updated velocities."""
    res = 24
    pos = _block(res, (7, 8, 7), (17, 14, 17), 4, 6)
    n = pos.shape[0]
    rng = np.random.default_rng(6)
    f = _fields(pos, 0.5 * rng.standard_normal((n, 3)), (1.0 / res) ** 3 / 4)
    f[5:8, :, :3] = 0.05 * rng.standard_normal((3, n, 3))
    f[2:8, :, 3] = 3.0 * rng.standard_normal((6, n))
    mats = mpm.materials_bytes([_material(1e4, 0.2, 1.0, mpm.MaterialModel.ELASTIC)])
    sc = oracle.make_scene(2e-3, 1.0 / res, (res, res, res), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), -1.0)
    gravity = mpm.GravityForce((0.0, 1.0, 0.0), -9.8)
    forces = [gravity, mpm.TurbulenceForce(strength=3.0, frequency=1.5, octaves=3, seed=0.37),
              mpm.NoiseForce((0.3, 0.8, 0.1), strength=4.0, frequency=2.0, octaves=2, seed=1.1)]
    sim = oracle.Sim(sc, mats)
    sim.set_state(f)
    sim.bin()
    sim.p2g()
    mom, mass = sim.get_grid()
    ids = np.nonzero(mass >= 1e-10)[0].astype(np.uint32)
    out = {}
    for key, fl in (("vel_forces", forces), ("vel_gravity", [gravity])):
        s2 = oracle.Sim(sc, mats, b"", mpm.forces_bytes(fl))
        s2.set_state(f)
        s2.set_grid(mom[: s2.NC], mass)
        s2.grid_update()
        m2, _ = s2.get_grid()
        out[key] = m2[ids]
    recs = np.frombuffer(mpm.forces_bytes(forces), np.uint8)
    return dict(res=np.asarray(res), dt=np.asarray(2e-3), forces=recs, node_ids=ids, mom_in=mom[ids],
                mass_in=mass[ids], **out)


def hash_case():
    """Unused slots (typeFlag 6) at the origin, as the reference's PBF engine
    leaves them, a clump of 300 particles in one cell, and particles spread
    over and beyond a non-cubic domain."""
    rng = np.random.default_rng(8)
    n = 3000
    p = np.zeros((n, 4), np.float32)
    tf = rng.integers(0, 6, n).astype(np.uint32)
    tf[:500] = 6
    p[500:800, :3] = rng.uniform(0.41, 0.43, (300, 3))
    p[800:, :3] = rng.uniform(-0.1, 1.1, (n - 800, 3)) * np.array([1.0, 0.5, 1.0], np.float32)
    p[:, 3] = rng.uniform(0.5, 2.0, n)
    perm = rng.permutation(n)
    p, tf = p[perm], tf[perm]
    wmin, h, gr = np.array([0.0, 0.0, 0.0], np.float32), np.float32(1.0 / 24), np.array([24, 12, 24], np.uint32)
    cnt, end, srt = oracle.hash_build(p, tf, wmin, float(h), tuple(int(x) for x in gr))
    return dict(pred=p, type_flag=tf, world_min=wmin, cell_size=np.asarray(h), grid_res=gr, cell_count=cnt,
                cell_end=end, sorted=srt)


def build_all():
    out = dict(substep_cases())
    out["bins_edge"] = bins_case()
    out["mesh_box_sdf"] = mesh_case()
    out["noise"] = noise_case()
    out["emit"] = emit_case()
    out["forces_grid"] = forces_grid_case()
    out["hash_pbf"] = hash_case()
    return out


def main():
    for name, arrs in build_all().items():
        path = os.path.join(HERE, name + ".npz")
        np.savez_compressed(path, **{k: np.asarray(v) for k, v in arrs.items()})
        print(f"{path}: {os.path.getsize(path)} bytes")


if __name__ == "__main__":
    main()
