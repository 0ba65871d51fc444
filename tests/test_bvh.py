"""BVH build (Util/BVH.fs:99-247): the product's iterative builder, the
oracle's recursive restatement and a third, pure-Python restatement (float32
scalars, small inputs) must produce identical node arrays and permutations;
structural invariants and the SAH split choice are checked directly."""
import numpy as np
import pytest

from barnacle_amd import configs, native, scene

F = np.float32
INF = F(np.inf)


def mn(a, b):  # Vector3.MinNative component (x86 minps)
    return a if a < b else b


def mx(a, b):
    return a if a > b else b


def union(A, B):
    return ([mn(A[0][i], B[0][i]) for i in range(3)], [mx(A[1][i], B[1][i]) for i in range(3)])


def union_pt(A, p):
    return ([mn(A[0][i], p[i]) for i in range(3)], [mx(A[1][i], p[i]) for i in range(3)])


def empty():
    return ([INF] * 3, [-INF] * 3)


def centroid(A):
    return [F(0.5) * (A[0][i] + A[1][i]) for i in range(3)]


def sa(A):
    d = [A[1][i] - A[0][i] for i in range(3)]
    with np.errstate(invalid="ignore", over="ignore"):
        return F(2) * (d[0] * d[1] + d[1] * d[2] + d[2] * d[0])


def py_build(boxes):
    """Tree.Build + Flatten + BVHNode.Build, restated in Python."""
    items = [((list(map(F, b[:3])), list(map(F, b[3:]))), i) for i, b in enumerate(boxes)]
    nodes = []

    def build(first, last, depth):
        n = last - first
        sub = items[first:last]
        bounds = empty()
        for b, _ in sub:
            bounds = union(bounds, b)
        me = len(nodes)
        nodes.append(None)
        if n <= 4 or depth >= 64:
            nodes[me] = (bounds, first, True, 0, n)
            return
        cb = empty()
        for b, _ in sub:
            cb = union_pt(cb, centroid(b))
        d = [cb[1][i] - cb[0][i] for i in range(3)]
        axis = 0 if (d[0] >= d[1] and d[0] >= d[2]) else (1 if d[1] >= d[2] else 2)
        ext = d[axis]
        if ext == 0:
            mid = first + n // 2
            build(first, mid, depth + 1)
            rc = len(nodes)
            build(mid, last, depth + 1)
            nodes[me] = (bounds, rc, False, axis, 0)
            return

        def bin_of(b):
            return min(11, int((F(12) * (centroid(b)[axis] - cb[0][axis])) / ext))
        bins = [[empty(), 0] for _ in range(12)]
        for b, _ in sub:
            k = bin_of(b)
            bins[k][0] = union(bins[k][0], b)
            bins[k][1] += 1
        costs = [F(0)] * 11
        lb, rb, lc, rcnt = empty(), empty(), 0, 0
        with np.errstate(invalid="ignore", over="ignore"):
            for i in range(11):
                lb = union(lb, bins[i][0])
                lc += bins[i][1]
                costs[i] = costs[i] + F(lc) * sa(lb)
                rb = union(rb, bins[11 - i][0])
                rcnt += bins[11 - i][1]
                costs[10 - i] = costs[10 - i] + F(rcnt) * sa(rb)
            best, bestc = 0, INF
            for i in range(11):
                c = costs[i] + F(n) * sa(cb)
                if c < bestc:
                    bestc, best = c, i
        left, right = first, last - 1
        for it in sub:
            if bin_of(it[0]) > best:
                items[right] = it
                right -= 1
            else:
                items[left] = it
                left += 1
        build(first, left, depth + 1)
        rc = len(nodes)
        build(right + 1, last, depth + 1)
        nodes[me] = (bounds, rc, False, axis, 0)

    build(0, len(items), 0)
    return nodes, [i for _, i in items]


def as_tuples(arr):
    return [([F(v) for v in r["pmin"]], [F(v) for v in r["pmax"]], int(r["right_or_offset"]), bool(r["is_leaf"]),
             int(r["split_axis"]), int(r["count"])) for r in arr]


def random_boxes(rng, n, grid=None):
    c = rng.uniform(-10, 10, (n, 3)) if grid is None else rng.randint(0, grid, (n, 3)).astype(float)
    e = rng.uniform(0, 2, (n, 3))
    return np.concatenate([c - e, c + e], axis=1).astype(np.float32)


CASES = [("uniform", 5, None), ("uniform", 17, None), ("uniform", 64, None),
         ("grid", 40, 3), ("grid", 33, 1), ("grid", 64, 2)]


@pytest.mark.parametrize("kind,n,grid", CASES)
def test_three_restatements_agree(orc, kind, n, grid):
    rng = np.random.RandomState(n * 31 + (grid or 0))
    b = random_boxes(rng, n, grid)
    pn, pp = native.bvh_build(b)
    on, op = orc.bvh_build(b)
    yn, yp = py_build(b.tolist())
    assert pn.tobytes() == on.tobytes()
    assert np.array_equal(pp, op)
    assert list(pp) == yp
    got = as_tuples(pn)
    assert len(got) == len(yn)
    for g, (bounds, word, leaf, axis, cnt) in zip(got, yn):
        assert g[0] == bounds[0] and g[1] == bounds[1]
        assert (g[2], g[3], g[5]) == (word, leaf, cnt)
        if not leaf:
            assert g[4] == axis


def check_invariants(nodes, perm, boxes):
    n_items = len(boxes)
    assert sorted(perm.tolist()) == list(range(n_items))
    covered = np.zeros(n_items, int)

    def rec(i, depth, parent_box):
        r = nodes[i]
        lo, hi = r["pmin"], r["pmax"]
        if parent_box is not None:
            assert np.all(lo >= parent_box[0]) and np.all(hi <= parent_box[1])
        if r["is_leaf"]:
            off, cnt = int(r["right_or_offset"]), int(r["count"])
            assert cnt <= 4 or depth >= 64
            pb = boxes[perm[off:off + cnt]]
            assert np.array_equal(lo, pb[:, :3].min(axis=0)) and np.array_equal(hi, pb[:, 3:].max(axis=0))
            covered[off:off + cnt] += 1
            return i + 1
        assert r["right_or_offset"] > i + 1
        end_left = rec(i + 1, depth + 1, (lo, hi))
        assert end_left == r["right_or_offset"]          # pre-order: right follows left subtree
        return rec(int(r["right_or_offset"]), depth + 1, (lo, hi))

    assert rec(0, 0, None) == len(nodes)
    assert np.all(covered == 1)


@pytest.mark.parametrize("n", [1, 2, 4, 5, 100, 5000])
def test_invariants_random(orc, n):
    rng = np.random.RandomState(n)
    b = random_boxes(rng, n)
    nodes, perm = native.bvh_build(b)
    on, op = orc.bvh_build(b)
    assert nodes.tobytes() == on.tobytes() and np.array_equal(perm, op)
    check_invariants(nodes, perm, b)


def test_sah_choice_is_minimal_at_root():
    """At the root the partition must realise the cheapest of the 11 binned
    SAH candidates (recomputed here in float64 from the resulting split)."""
    rng = np.random.RandomState(7)
    b = random_boxes(rng, 300)
    nodes, perm = native.bvh_build(b)
    c = 0.5 * (b[:, :3] + b[:, 3:])
    cmin, cmax = c.min(axis=0), c.max(axis=0)
    axis = int(np.argmax(cmax - cmin))
    binidx = np.minimum(11, ((12 * (c[:, axis] - cmin[axis])) / (cmax[axis] - cmin[axis])).astype(int))

    def cost(mask):
        def area(bb):
            d = bb[:, 3:].max(axis=0) - bb[:, :3].min(axis=0)
            return 2 * (d[0] * d[1] + d[1] * d[2] + d[2] * d[0])
        return mask.sum() * area(b[mask]) + (~mask).sum() * area(b[~mask])
    costs = [cost(binidx <= s) for s in range(11) if 0 < (binidx <= s).sum() < len(b)]
    left_n = int(nodes[int(nodes[0]["right_or_offset"])]["right_or_offset"])  # unused guard
    rc = int(nodes[0]["right_or_offset"])
    # the left subtree covers perm[0:k]; find k as the number of items in bins <= best
    n_left = 0
    stack = [1]
    while stack:
        i = stack.pop()
        r = nodes[i]
        if r["is_leaf"]:
            n_left += int(r["count"])
        else:
            stack += [i + 1, int(r["right_or_offset"])]
    chosen = cost(np.isin(np.arange(len(b)), perm[:n_left]))
    assert chosen <= min(costs) * (1 + 1e-5)
    del left_n, rc


def test_blas_of_config_mesh_matches_oracle(orc):
    sc = scene.load_scene(configs.config1(width=16, height=16, spp=1, freq=20))
    osc = orc.scene(sc)
    p = len(sc.primitives) - 1
    ob = osc.export_blas(p)
    V = sc.vertices.reshape(-1, 3)[sc.primitives[p]["vert_offset"]:][:sc.primitives[p]["n_verts"]]
    I = sc.indices[sc.primitives[p]["index_offset"]:][:sc.primitives[p]["n_indices"]].reshape(-1, 3)
    T = V[I]
    nodes, perm = native.bvh_build(np.concatenate([T.min(axis=1), T.max(axis=1)], axis=1))
    assert nodes.tobytes() == ob["nodes"].tobytes()
    assert np.array_equal(I[perm].reshape(-1), ob["tri"])
    check_invariants(nodes, perm, np.concatenate([T.min(axis=1), T.max(axis=1)], axis=1))


def test_negative_bin_index_raises_instead_of_writing_out_of_bounds():
    """Non-finite item bounds can make the binned-SAH index negative (a NaN
    centroid leaves the running minimum above other centroids); the
    reference's bounds-checked Span throws there (Util/BVH.fs:163-174), and
    the host build returns an error instead of writing outside its bins."""
    import pytest
    from barnacle_amd import native
    b = np.array([[0, 0, 0, 1, 1, 1], [2, 0, 0, 3, 1, 1], [4, 0, 0, 5, 1, 1], [np.nan] * 6,
                  [6, 0, 0, 7, 1, 1], [8, 0, 0, 9, 1, 1], [10, 0, 0, 11, 1, 1]], np.float32)
    with pytest.raises(native.BarnacleError, match="bin index out of range"):
        native.bvh_build(b)
