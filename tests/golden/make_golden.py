"""Writes tests/golden/ref_vectors.npz from the reference's own code.

The vectors come from oracle/_ref/libvflipref.so, which is the reference's
levelset_util.cpp, fluid_buffer3D.h and util.h compiled as-is (see
oracle/Makefile, target `ref`). Only inputs and outputs are stored.

Run from the repository root:  make -C oracle ref && python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402


def fraction_cases(rng):
    # random corners, plus exact zeros / signs / ties that exercise every branch
    r = rng.standard_normal((4000, 4)).astype(np.float32)
    special = np.array([[-1, -1, -1, -1], [1, 1, 1, 1], [-1, 1, 1, 1], [1, -1, 1, 1], [1, 1, -1, 1],
                        [1, 1, 1, -1], [-1, -1, 1, 1], [-1, 1, -1, 1], [-1, 1, 1, -1], [1, -1, -1, 1],
                        [-0.5, 0.25, 0.25, -0.5], [-0.5, 0.6, 0.7, -0.4], [0.5, -0.6, -0.7, 0.4],
                        [-0.1, 0.1, 0.1, -0.1], [0, -1, -1, -1], [-1, 0, -1, -1], [0, 1, -1, 1],
                        [-2, 3, -0.5, 1e-7], [1e-30, -1e-30, 1e-30, -1e-30]], np.float32)
    q = rng.integers(-2, 3, (2000, 4)).astype(np.float32) * np.float32(0.5)
    # avoid corner sets on which the reference's rotation loop never ends
    # (the loop needs a strictly negative anchor; inside counts phi <= 0)
    allc = np.concatenate([r, special, q])
    keep = []
    for c in allc:
        inside = int((c <= 0).sum())
        neg = int((c < 0).sum())
        if inside in (1, 2) and neg == 0:
            continue
        if inside == 2 and neg == 1:
            continue
        keep.append(c)
    return np.array(keep, np.float32)


def main():
    R = O.ref_lib()
    if R is None:
        raise SystemExit("oracle/_ref/libvflipref.so missing: make -C oracle ref")
    rng = np.random.default_rng(20240601)
    out = {}
    fc = fraction_cases(rng)
    out["frac4_in"] = fc
    out["frac4_out"] = np.array([R.ref_fraction_inside4(*map(float, c)) for c in fc], np.float32)
    f2 = rng.standard_normal((500, 2)).astype(np.float32)
    f2 = np.concatenate([f2, np.array([[0, 0], [-1, 0], [0, -1], [1, -1], [-1, 1], [-1, -1]], np.float32)])
    out["frac2_in"] = f2
    out["frac2_out"] = np.array([R.ref_fraction_inside2(*map(float, c)) for c in f2], np.float32)
    seeds = np.concatenate([np.arange(64), rng.integers(0, 2**32 - 1, 200)]).astype(np.uint32)
    out["randhash_in"] = seeds
    out["randhash_out"] = np.array([R.ref_randhash(int(s)) for s in seeds], np.uint32)
    # samplers: an odd-sized field, offsets of a u-face grid, positions inside and outside
    nx, ny, nz = 11, 9, 7
    h = 0.125
    field = rng.standard_normal((nz, ny, nx)).astype(np.float32)
    pts = np.concatenate([rng.uniform(-0.3, 1.6, (400, 3)), rng.uniform(0, 1, (100, 3))]).astype(np.float64)
    out["samp_field"] = field
    out["samp_meta"] = np.array([nx, ny, nz, h, 0.5, 0.0, 0.0], np.float64)
    out["samp_pts"] = pts
    for kind, name in ((0, "linear"), (1, "cubic"), (2, "cube_lerp"), (3, "grad")):
        o = np.zeros(len(pts), np.float32)
        c8 = np.zeros(8 * len(pts), np.float32)
        o3 = np.zeros(3 * len(pts), np.float32)
        R.ref_sample(kind, np.ascontiguousarray(field), nx, ny, nz, h, 0.5, 0.0, 0.0, len(pts),
                     np.ascontiguousarray(pts), o, c8, o3)
        out[f"samp_{name}"] = o3.reshape(-1, 3) if kind == 3 else o
        if kind == 2:
            out["samp_corners"] = c8.reshape(-1, 8)
    # blocked storage index
    bi = []
    for (a, b, c) in ((11, 9, 7), (16, 16, 8), (17, 3, 9)):
        for (i, j, k) in ((0, 0, 0), (a - 1, b - 1, c - 1), (a // 2, b // 3, c // 2), (8, 1, 2)):
            if i < a and j < b and k < c:
                bi.append((a, b, c, i, j, k, R.ref_blocked_index(a, b, c, i, j, k)))
    out["blocked_index"] = np.array(bi, np.int64)
    # normalize(Vec3f): random, tiny, huge and zero vectors
    nv = np.concatenate([rng.standard_normal((2000, 3)), rng.standard_normal((100, 3)) * 1e-7,
                         rng.standard_normal((100, 3)) * 1e6, np.zeros((1, 3))]).astype(np.float32)
    no = np.zeros_like(nv)
    R.ref_normalize(len(nv), np.ascontiguousarray(nv), no)
    out["normalize_in"] = nv
    out["normalize_out"] = no
    # hash_FLIP cell indices: positions around cell faces, negatives in (-h, 0),
    # far outside, for h = 1/64 and 1/48, coarse and subdomain (bmin) forms
    ci_pos, ci_meta, ci_out = [], [], []
    for h, sub, bmin in ((1.0 / 64, 0, (0.0, 0.0, 0.0)), (1.0 / 48, 1, (0.3, 0.25, 0.4))):
        hf = np.float32(h)
        base = rng.integers(-3, 70, (3000, 3)).astype(np.float32)
        frac = rng.choice(np.array([-0.5, -0.5000001, -0.4999999, 0.0, 0.4999999, 0.5, rng.uniform(-1, 1)],
                                   np.float32), (3000, 3))
        p = ((base + frac) * hf).astype(np.float32) + np.float32(bmin[0] if sub else 0)
        p = np.concatenate([p, rng.uniform(-0.02, 1.05, (3000, 3)).astype(np.float32)])
        o = np.zeros(p.shape, np.int32)
        R.ref_cell_index(len(p), np.ascontiguousarray(p), float(hf), sub, *bmin, o)
        ci_pos.append(p)
        ci_out.append(o)
        ci_meta.append([float(hf), sub, *bmin])
    out["cellidx_pos"] = np.stack(ci_pos)
    out["cellidx_out"] = np.stack(ci_out)
    out["cellidx_meta"] = np.array(ci_meta, np.float64)
    # set_boundary_phi / set_boundary_vel sample positions: nodal grid and the three
    # face grids, with their Buffer3D offsets, for two spacings
    gp = []
    for h in (1.0 / 64, 1.0 / 48):
        hf = np.float32(h)
        for offs, shp in (((0.5, 0.5, 0.5), (9, 10, 11)), ((0.5, 0.0, 0.0), (8, 9, 11)), ((0.0, 0.5, 0.0), (8, 10, 10)),
                          ((0.0, 0.0, 0.5), (9, 9, 10))):
            nz, ny, nx = shp
            o = np.zeros((nz, ny, nx, 3), np.float32)
            R.ref_grid_points(nx, ny, nz, float(hf), *offs, o)
            gp.append((float(hf), offs, shp, o))
    out["gridpts_meta"] = np.array([[g[0], *g[1], *g[2]] for g in gp], np.float64)
    out["gridpts_out"] = np.concatenate([g[3].reshape(-1) for g in gp])
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ref_vectors.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, {k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    main()
