"""The host mesh deformation (fsi_deform_*) against a dense numpy restatement of
LaplacianDeformation (src/deformation/LaplacianDeformation.cpp) and deformStep
(src/main.cpp:862-944).

The restatement builds the cotangent Laplacian densely from its definition,
pinned here by closed forms on a square-grid plate (an axis edge between two
45-degree corners weighs 1, a diagonal opposite two right angles takes the 1e-4
floor), and solves with numpy's dense LU or least squares instead of the
library's sparse LDL^T.  The reference's mesh keeps its points as floats, so
the library rounds the rest pose and every solved position to float: the
plates here have float vertices, and a solve is compared with the dense
solution rounded to float, to one float ulp (the two FP64 solutions differ
far below it, but may straddle a rounding boundary)."""
import numpy as np
import pytest

import fsi_pbf


def as_mesh(x):
    """Positions as the reference's mesh stores them: rounded to float."""
    return np.asarray(x, dtype=np.float64).astype(np.float32).astype(np.float64)


def assert_mesh_close(got, want, extra=0.0):
    """got holds floats and is the FP64 solution want rounded to float, to one
    float ulp (plus `extra` for a restatement that differs more)."""
    assert np.array_equal(got, as_mesh(got)), "vertices must be float values"
    ulp = np.spacing(np.abs(want).astype(np.float32)).astype(np.float64)
    err = np.abs(got - want) - ulp - extra
    assert err.max() <= 1e-13, float(err.max())


def plate(nx=12, ny=6, size=(2.4, 1.2)):
    xs = np.linspace(0.0, size[0], nx + 1)
    ys = np.linspace(0.0, size[1], ny + 1)
    X, Y = np.meshgrid(xs, ys, indexing="xy")
    verts = as_mesh(np.stack([X.ravel(), Y.ravel(), np.zeros(X.size)], axis=1))
    faces = []
    for j in range(ny):
        for i in range(nx):
            a = j * (nx + 1) + i
            b, c, d = a + 1, a + nx + 1, a + nx + 2
            faces += [[a, b, d], [a, d, c]]
    return verts, np.array(faces, dtype=np.int32)


def cot_weights(verts, faces):
    n = len(verts)
    acc = {}
    for f in faces:
        for e in range(3):
            s, t, o = int(f[e]), int(f[(e + 1) % 3]), int(f[(e + 2) % 3])
            a, b = verts[s] - verts[o], verts[t] - verts[o]
            sn = np.linalg.norm(np.cross(a, b))
            key = (min(s, t), max(s, t))
            acc.setdefault(key, 0.0)
            if sn > 1e-6:
                acc[key] += np.dot(a, b) / sn
    W = np.zeros((n, n))
    for (i, j), c in acc.items():
        W[i, j] = W[j, i] = max(0.5 * c, 1e-4)
    return W


def laplacian(W):
    L = np.eye(len(W))
    rs = W.sum(axis=1)
    nz = rs > 0
    L[nz] -= W[nz] / rs[nz, None]
    return L


def test_weights_closed_form():
    verts, faces = plate(4, 4, (1.0, 1.0))
    W = cot_weights(verts, faces)
    # interior axis edge 6-7 (row 1): two opposite 45-degree corners
    assert W[6, 7] == pytest.approx(1.0, abs=1e-12)
    # boundary axis edge 0-1: one face
    assert W[0, 1] == pytest.approx(0.5, abs=1e-12)
    # diagonal 0-6: opposite two right angles
    assert W[0, 6] == 1e-4
    L = laplacian(W)
    assert np.allclose(L.sum(axis=1), 0.0, atol=1e-12)


def kring(W, seeds, k):
    seen = list(dict.fromkeys(seeds))
    front = list(seen)
    for _ in range(k):
        nxt = []
        for v in front:
            for u in np.nonzero(W[v])[0]:
                if u not in seen:
                    seen.append(int(u))
                    nxt.append(int(u))
        front = nxt
    return seen


def local_solve(cur, rest, W, anchors, handles, targets, k=3):
    """precomputeLocalSystem + solveLocal, dense."""
    L = laplacian(W)
    delta = L @ rest
    ring = kring(W, list(handles), k)
    interior = sorted(v for v in ring if v not in anchors)
    iset = set(interior)
    out = cur.copy()
    ni = len(interior)
    Lii = np.zeros((ni, ni))
    bnd = sorted({int(u) for v in interior for u in np.nonzero(W[v])[0] if int(u) not in iset} | {v for v in ring if v in anchors})
    Lib = np.zeros((ni, len(bnd)))
    for li, gi in enumerate(interior):
        ws = max(W[gi].sum(), 1e-12)
        Lii[li, li] = 1.0
        for gj in np.nonzero(W[gi])[0]:
            w = W[gi, gj] / ws
            if int(gj) in iset:
                Lii[li, interior.index(int(gj))] = -w
            else:
                Lib[li, bnd.index(int(gj))] = -w
    A = Lii.T @ Lii
    rhs = Lii.T @ (delta[interior] - Lib @ cur[bnd])
    for h, t in zip(handles, targets):
        if h in iset:
            A[interior.index(h), interior.index(h)] += 1.0
            rhs[interior.index(h)] += t
    out[interior] = np.linalg.solve(A, rhs)
    return out


def bent_targets(verts, handles, amp):
    return verts[handles] + np.outer(amp * np.sin(np.arange(len(handles)) + 1.0), [0.0, 0.0, 1.0])


@pytest.mark.parametrize("k", [1, 3])
def test_local_patch_matches_the_dense_restatement(k):
    verts, faces = plate()
    W = cot_weights(verts, faces)
    anchors = list(range(0, 13 * 7, 13))  # the x = 0 column
    handles = [45, 46, 58, 85]
    targets = bent_targets(verts, handles, 0.08)
    with fsi_pbf.Deform(verts, faces, kring=k) as d:
        for a in anchors:
            d.add_anchor(a)
        d.set_control_points(handles, targets)
        d.precompute()
        d.solve()
        got = d.vertices()
        patch = d.info()["patch"]
    want = local_solve(verts, verts, W, set(anchors), handles, targets, k)
    assert patch == len([v for v in kring(W, handles, k) if v not in anchors])
    assert_mesh_close(got, want)
    assert np.array_equal(got[[v for v in range(len(verts)) if v not in kring(W, handles, k)]],
                          verts[[v for v in range(len(verts)) if v not in kring(W, handles, k)]])


def test_handles_at_rest_keep_the_mesh():
    verts, faces = plate()
    with fsi_pbf.Deform(verts, faces) as d:
        d.add_anchor(0)
        d.set_control_points([40, 41], verts[[40, 41]])
        d.precompute()
        d.solve()
        got = d.vertices()
        assert np.array_equal(got, as_mesh(got)) and np.abs(got - verts).max() < 1e-12


def test_global_soft_path_is_least_squares():
    verts, faces = plate(8, 4)
    W = cot_weights(verts, faces)
    L = laplacian(W)
    anchors = [0, 9, 18]
    handles = [30, 40]
    targets = bent_targets(verts, handles, 0.1)
    with fsi_pbf.Deform(verts, faces, use_local=False) as d:
        for a in anchors:
            d.add_anchor(a)
        d.set_control_points(handles, targets)
        d.precompute()
        d.solve()
        got = d.vertices()
    n = len(verts)
    C = np.zeros((len(anchors) + len(handles), n))
    for r, v in enumerate(anchors + handles):
        C[r, v] = 1.0
    A = np.vstack([L, C])
    want = np.stack([np.linalg.lstsq(A, np.concatenate([L @ verts[:, a], verts[anchors, a], targets[:, a]]), rcond=None)[0]
                     for a in range(3)], axis=1)
    assert_mesh_close(got, want, 1e-9)


def hard_solve(verts, W, constrained, targets_by_vertex):
    """Eq. 14 system of the incremental path, dense."""
    L = laplacian(W)
    A = L.T @ L + 1e-8 * np.eye(len(verts))
    out = np.zeros_like(verts)
    c = sorted(constrained)
    free = [v for v in range(len(verts)) if v not in constrained]
    for a in range(3):
        b = L.T @ (L @ verts[:, a])
        t = np.zeros(len(verts))
        for v in c:
            t[v] = targets_by_vertex[v][a]
        rhs = b - A[:, c] @ t[c]
        x = np.zeros(len(verts))
        x[c] = t[c]
        x[free] = np.linalg.solve(A[np.ix_(free, free)], rhs[free])
        out[:, a] = x
    return out


def test_incremental_path_hard_constraints_and_updates():
    verts, faces = plate(10, 5)
    W = cot_weights(verts, faces)
    anchors = list(range(0, 11 * 6, 11))
    with fsi_pbf.Deform(verts, faces, use_local=False, incremental=True) as d:
        for a in anchors:
            d.add_anchor(a)
        rng = np.random.default_rng(5)
        steps = []
        for it in range(4):
            handles = sorted(rng.choice([v for v in range(len(verts)) if v not in anchors], size=3 + it, replace=False).tolist())
            targets = bent_targets(verts, handles, 0.05 * (it + 1))
            d.set_control_points(handles, targets)
            d.precompute()
            d.solve()
            got = d.vertices()
            info = d.info()
            steps.append(info["factor_steps"])
            tv = {v: verts[v] for v in anchors}
            tv.update({h: t for h, t in zip(handles, targets)})
            want = hard_solve(verts, W, set(anchors) | set(handles), tv)
            assert_mesh_close(got, want, 1e-8)
            assert np.array_equal(got[handles], as_mesh(targets))  # hard: the targets are hit, as floats
            assert d.verify_incremental() < 1e-9
        assert steps[0] == len(verts)
    # a local change of the constrained set recomputes only the tree paths
    # from the changed columns (fewer steps than a full factorization)
    verts, faces = plate(40, 10, (4.0, 1.0))
    with fsi_pbf.Deform(verts, faces, use_local=False, incremental=True) as d:
        for a in range(0, 41 * 11, 41):
            d.add_anchor(a)
        d.set_control_points([200], verts[[200]] + [0.0, 0.0, 0.05])
        d.precompute()
        assert d.info()["factor_steps"] == len(verts)
        d.set_control_points([200, 439], verts[[200, 439]] + [0.0, 0.0, 0.05])
        d.precompute()
        d.solve()
        assert 0 < d.info()["factor_steps"] < len(verts)
        assert d.verify_incremental() < 1e-9


def step_restated(cur, rest, forces, dt, last, anchors, p):
    """deformStep's handle rule (src/main.cpp:862-914)."""
    f32 = np.float32
    hs, ts = [], []
    for i in range(len(cur)):
        if i in anchors:
            continue
        f = forces[i].astype(np.float32)
        fmag = f32(np.sqrt(f32(f32(f[0] * f[0] + f[1] * f[1]) + f[2] * f[2])))
        was = i in last
        pushed = fmag >= f32(p.control_threshold) or (was and fmag >= f32(p.control_threshold_low))
        relaxing = bool(p.restore_enabled) and was and np.linalg.norm(cur[i] - rest[i]) > 1e-3
        if not (pushed or relaxing):
            continue
        t = cur[i].copy()
        if pushed:
            g = f32(f32(p.stiffness) * f32(dt))
            disp = (f * g).astype(np.float32)
            dm = f32(np.sqrt(f32(f32(disp[0] * disp[0] + disp[1] * disp[1]) + disp[2] * disp[2])))
            if dm > f32(p.max_displacement) and dm > f32(1e-12):
                disp = (disp * f32(f32(p.max_displacement) / dm)).astype(np.float32)
            t = t + disp.astype(np.float64)
        if p.restore_enabled:
            t = t + float(f32(p.restore_strength)) * (rest[i] - t)
        fr = t - rest[i]
        off = np.linalg.norm(fr)
        if off > float(f32(p.max_total_disp)):
            t = rest[i] + fr * (float(f32(p.max_total_disp)) / off)
        hs.append(i)
        ts.append(t)
    return hs, np.array(ts).reshape(-1, 3)


def test_deform_step_follows_the_reference_rule():
    verts, faces = plate()
    W = cot_weights(verts, faces)
    anchors = set(range(0, 13 * 7, 13))
    p = fsi_pbf.deform_params()
    rng = np.random.default_rng(11)
    cur = verts.copy()
    last = []
    with fsi_pbf.Deform(verts, faces) as d:
        for a in sorted(anchors):
            d.add_anchor(a)
        for frame in range(6):
            forces = np.zeros((len(verts), 3), dtype=np.float32)
            if frame < 3:  # a pressure patch, then nothing (the mesh relaxes)
                hit = rng.choice(len(verts), size=10, replace=False)
                forces[hit, 2] = rng.uniform(0.2, 3.0, size=10).astype(np.float32)
            nh, solved = d.step(forces, 1.0 / 120.0, p)
            hs, ts = step_restated(cur, verts, forces, 1.0 / 120.0, last, anchors, p)
            assert nh == len(hs)
            if hs or hs != last:
                assert solved
                if hs:
                    want = local_solve(cur, verts, W, anchors, hs, ts)
                    assert_mesh_close(d.vertices(), want, 1e-6 * frame)
                    cur = as_mesh(want)
            last = hs
            assert np.abs(d.vertices() - cur).max() < 1e-6
        assert d.info()["factorizations"] >= 2


def test_vertices_are_the_reference_mesh_floats():
    """LaplacianDeformation writes solved positions into an OpenMesh mesh of
    float points (LaplacianDeformation.cpp:388-389), and its rest pose comes
    from that mesh: vertices given in FP64 come back rounded to float, and a
    solve moves every patch vertex to the float nearest its FP64 solution."""
    xs = np.linspace(0.0, 1.0, 11)  # 0.1 steps: not floats
    X, Y = np.meshgrid(xs, xs[:6], indexing="xy")
    verts = np.stack([X.ravel(), Y.ravel(), 0.1 * X.ravel()], axis=1)
    faces = plate(10, 5, (1.0, 0.5))[1]
    assert not np.array_equal(verts, as_mesh(verts))
    with fsi_pbf.Deform(verts, faces, kring=2) as d:
        assert np.array_equal(d.vertices(), as_mesh(verts))
        d.add_anchor(0)
        targets = as_mesh(verts)[[27, 28]] + [[0.0, 0.0, 0.013], [0.0, 0.0, -0.007]]
        d.set_control_points([27, 28], targets)
        d.precompute()
        d.solve()
        got = d.vertices()
    mv = as_mesh(verts)
    want = local_solve(mv, mv, cot_weights(mv, faces), {0}, [27, 28], targets, 2)
    moved = np.any(got != mv, axis=1)
    assert moved.sum() > 2
    assert_mesh_close(got, want)
