"""The oracle's uniform-grid neighbour search (UGNSCudaImpl.cu:28-227)
against an independent NumPy restatement: cell ids from FP32 floor((x-LB)/h),
a stable sort on (cell id, attach id) (np.lexsort), cell ranges by scan, and
neighbour sets by O(N^2) brute force with the same FP32 distance test,
listed in the reference traversal order (self first, then the 27-cell offset
table with dx outermost, members in sorted order), truncated at M.  Plus the
exact-lattice count: 33 neighbours (self included) per interior particle.
"""
import json

import numpy as np
import pytest

import pyoracle
import vtp
from vtp import scenes

f32 = np.float32


def _params(lb, size, r=0.01, M=60):
    cfg = scenes.pbf_config(radius=r, lb=lb, size=size, max_nbr=M)
    return vtp.params_from_json("PBF", json.dumps(cfg)), cfg


def _grid(lb, size, h):
    return [int(np.ceil(f32(size[a]) / f32(h))) for a in range(3)]


def _numpy_ugns(pos, lb, size, h, M):
    pos = pos.astype(np.float32)
    n = pos.shape[0]
    g = _grid(lb, size, h)
    cp = np.floor((pos - np.array(lb, np.float32)) / np.float32(h)).astype(np.int64)
    alive = np.all((cp >= 0) & (cp < np.array(g)), axis=1) & np.all(np.isfinite(pos), axis=1)
    C = g[0] * g[1] * g[2]
    key = np.where(alive, cp[:, 2] * g[1] * g[0] + cp[:, 1] * g[0] + cp[:, 0], C)
    order = np.lexsort((np.arange(n), key))          # stable on (key, attach id)
    skey = key[order]
    start = np.full(C + 1, 0xFFFFFFFF, np.uint64)
    end = np.full(C + 1, 0xFFFFFFFF, np.uint64)
    for i in range(n):
        k = skey[i]
        if i == 0 or skey[i - 1] != k:
            start[k] = i
        if i == n - 1 or skey[i + 1] != k:
            end[k] = i
    # FP32 distance test with every operation rounded: ((dx*dx + dy*dy) + dz*dz)
    d = pos[:, None, :] - pos[None, :, :]
    d2 = (d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2]
    ln = np.sqrt(d2)
    adj = (ln > 0) & (ln <= np.float32(h))
    lists = np.full((n, M), 0xFFFFFFFF, np.uint64)
    num = np.zeros(n, np.uint64)
    offs = [(a, b, c) for a in (-1, 0, 1) for b in (-1, 0, 1) for c in (-1, 0, 1)]
    for p in range(n):
        if not alive[p]:
            continue
        row = [p]
        for ox, oy, oz in offs:
            X, Y, Z = cp[p, 0] + ox, cp[p, 1] + oy, cp[p, 2] + oz
            if min(X, Y, Z) < 0 or X >= g[0] or Y >= g[1] or Z >= g[2]:
                continue
            cid = Z * g[1] * g[0] + Y * g[0] + X
            if start[cid] == 0xFFFFFFFF:
                continue
            for j in range(int(start[cid]), int(end[cid]) + 1):
                q = order[j]
                if q != p and alive[q] and adj[p, q] and len(row) < M:
                    row.append(q)
        lists[p, :len(row)] = row
        num[p] = len(row)
    return order, start[:C], end[:C], lists, num, C


def _oracle(params, pos):
    n = pos.shape[0]
    o = pyoracle.Oracle(params, pos, None, None, None, (np.array([0], np.uint32), np.array([n], np.uint32),
                                                         np.array([vtp.MAT_FLUID], np.int32)))
    o.initialize()
    return o


@pytest.mark.parametrize("seed,M", [(0, 60), (1, 60), (2, 12)])
def test_oracle_ugns_matches_bruteforce(seed, M):
    lb, size = (-0.1, -0.05, 0.0), (0.37, 0.29, 0.33)
    params, _ = _params(lb, size, M=M)
    rng = np.random.default_rng(seed)
    n = 1500
    pos = rng.uniform([-0.1, -0.05, 0.0], [0.27, 0.24, 0.33], size=(n, 3)).astype(np.float32)
    pos[10:20] = pos[0]                                  # duplicates: distance 0 is not a neighbour
    pos[20:25] = [[-0.2, 0.0, 0.1]] * 5                  # below the grid: dead, empty lists
    pos[25:30] = np.array([[0.05, 0.05, 0.05]]) + np.arange(5)[:, None] * np.float32(0.04)  # exactly h apart
    o = _oracle(params, pos)
    order, cs, ce, lists, num, C = _numpy_ugns(pos, lb, size, 0.04, M)
    got = o.neighbors(C)
    assert np.array_equal(got["sorted"], order.astype(np.uint32))
    assert np.array_equal(got["cell_start"], cs.astype(np.uint32))
    assert np.array_equal(got["cell_end"], ce.astype(np.uint32))
    assert np.array_equal(got["neighbor_num"], num.astype(np.uint32))
    assert np.array_equal(got["neighbors"], lists.astype(np.uint32))
    if M == 12:
        assert (num == M).sum() > 100                   # the cap is exercised
    assert np.all(num[20:25] == 0)


def test_oracle_neighbour_sets_are_symmetric():
    lb, size = (0.0, 0.0, 0.0), (0.4, 0.4, 0.4)
    params, _ = _params(lb, size)
    rng = np.random.default_rng(7)
    pos = rng.uniform(0.01, 0.39, size=(2000, 3)).astype(np.float32)
    o = _oracle(params, pos)
    got = o.neighbors(_grid(lb, size, 0.04)[0] ** 3)
    L, num = got["neighbors"], got["neighbor_num"]
    assert num.max() < 60
    pairs = {(int(p), int(L[p, t])) for p in range(len(pos)) for t in range(1, int(num[p]))}
    assert all((q, p) in pairs for p, q in pairs)


def test_exact_lattice_interior_count_is_33():
    """Lattice spacing s = 2r, h = 2s: interior neighbourhoods are the 33
    integer offsets with i^2 + j^2 + k^2 <= 4 (self included)."""
    r = 0.01
    lb, size = (0.0, 0.0, 0.0), (0.4, 0.4, 0.4)
    params, _ = _params(lb, size, r=r)
    pos = scenes.particle_cube(r, (0.05, 0.05, 0.05), (0.2, 0.2, 0.2))
    o = _oracle(params, pos)
    got = o.neighbors(_grid(lb, size, 4 * r)[0] ** 3)
    num = got["neighbor_num"].astype(int)
    c = pos.mean(0)
    interior = np.all(np.abs(pos - c) < 0.2 / 2 - 2 * 0.02 - 1e-6, axis=1)
    assert interior.sum() > 100
    assert np.all(num[interior] == 33)
    assert sum(1 for i in range(-2, 3) for j in range(-2, 3) for k in range(-2, 3) if i * i + j * j + k * k <= 4) == 33
