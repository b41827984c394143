"""Ray-traced wall sampling (host code, no GPU): the product's traversal
(nodals_ray_trace, csrc/host/ray_trace.cpp) against the brute-force
checker oracle/ray_oracle.py, and the ray_trace placement mode of the
k-omega setup against the adjacency walk (legacy_endpoint).

Reference behaviour: turbulence/sst_gpu_v2/source/puretet_profile5_eq5_hxt5b/
modules/wall/portable/tet_ray_trace.hpp (strict containment: a segment that
reaches the boundary before its sample fails even when the sample point lies
in another part of the fluid, :2-5) and modules/wall/tet_wall_ray_cache.inc
(active slots, local shortening: 28 bisection steps, 0.90 safety, 0.25
floor, :118-152)."""
import numpy as np
import pytest

from nodals_amd import _lib
from oracle import ray_oracle as R


def _rays_from_boundary(pts, tets, rng, count):
    faces = R.boundary_faces(tets)
    pick = rng.choice(len(faces), size=count, replace=False)
    start, origin, direction = [], [], []
    for f in pick:
        c, j = faces[f]
        v, n = R.face_frame(pts, tets, c, j)
        w = rng.dirichlet([2.0, 2.0, 2.0])
        start.append(c)
        origin.append(w @ v)
        direction.append(n)
    return np.array(start), np.array(origin), np.array(direction)


def _check_hits(pts, tets, origin, direction, dist, out, tol=1e-9):
    for r in range(len(origin)):
        d = direction[r] / np.linalg.norm(direction[r])
        for m, t in enumerate(dist[r]):
            p = origin[r] + t * d
            c = out["cell"][r, m]
            assert c in R.locate_all(pts, tets, p, tol=10 * tol), (r, m)
            lam = out["bary"][r, m]
            assert abs(lam.sum() - 1) < 1e-12 and lam.min() >= -tol
            rec = lam @ pts[tets[c]]
            assert np.linalg.norm(rec - p) <= 1e-12 * max(1.0, np.linalg.norm(p))
        assert np.all(np.diff(out["crossings"][r]) >= 0)


@pytest.mark.parametrize("jitter", [0.0, 0.2])
def test_rays_inside_a_box_find_the_brute_force_cells(jitter):
    """Axis-aligned rays of the structured Kuhn box pass exactly through
    edges and vertices (the face-fan search); the jittered box is generic."""
    pts, tets = R.kuhn_box(5, 4, 4, 0.25, jitter=jitter, seed=3)
    rng = np.random.default_rng(11)
    start, origin, direction = _rays_from_boundary(pts, tets, rng, 40)
    # samples at fractions of the distance at which each ray leaves the box
    exits = np.array([R.segment_cover(pts, tets, origin[r], direction[r], 10.0)[1] for r in range(len(start))])
    dist = exits[:, None] * np.array([0.1, 0.35, 0.6, 0.85])[None, :]
    out = _lib.ray_trace(pts, tets, start, origin, direction, dist)
    assert np.all(out["status"] == 0)
    _check_hits(pts, tets, origin, direction, dist, out)
    assert out["crossings"].max() > 0


def test_rays_that_leave_the_box_fail_at_the_brute_force_exit():
    pts, tets = R.kuhn_box(4, 4, 3, 0.25, jitter=0.15, seed=5)
    rng = np.random.default_rng(2)
    start, origin, direction = _rays_from_boundary(pts, tets, rng, 25)
    exits = np.array([R.segment_cover(pts, tets, origin[r], direction[r], 10.0)[1] for r in range(len(start))])
    dist = np.stack([0.5 * exits, 1.5 * exits], axis=1)
    out = _lib.ray_trace(pts, tets, start, origin, direction, dist)
    assert np.all(out["status"] == 1)
    # certified length of the first sample = the exit distance, when the first
    # sample is past it
    out2 = _lib.ray_trace(pts, tets, start, origin, direction, (1.5 * exits)[:, None])
    assert np.all(out2["status"] == 1)
    assert np.allclose(out2["certified"], exits, rtol=0, atol=1.5 * exits.max() * 2.0 ** -27 + 1e-9)


def test_segment_leaving_the_fluid_is_rejected_even_where_the_sample_is_fluid():
    """U channel: from the left arm's outer wall straight across the notch
    into the right arm.  The sample point is inside the mesh (brute force
    finds its cell), but the segment crosses the notch, so the trace fails;
    its certified length is the left arm's width."""
    pts, tets = R.u_channel(3, 1.0 / 3)
    faces = [(c, j) for c, j in R.boundary_faces(tets)
             if np.allclose(R.face_frame(pts, tets, c, j)[0][:, 0], 0.0)
             and R.face_frame(pts, tets, c, j)[0][:, 1].min() >= 1.0 - 1e-12]
    assert faces
    start, origin, direction = [], [], []
    for c, j in faces:
        v, n = R.face_frame(pts, tets, c, j)
        start.append(c)
        origin.append(v.mean(axis=0))
        direction.append(n)
    start, origin, direction = np.array(start), np.array(origin), np.array(direction)
    assert np.allclose(direction, [1, 0, 0])
    for r in range(len(start)):
        assert len(R.locate_all(pts, tets, origin[r] + 2.5 * direction[r])) > 0
        assert not R.segment_cover(pts, tets, origin[r], direction[r], 2.5)[0]
    out = _lib.ray_trace(pts, tets, start, origin, direction, np.full((len(start), 1), 2.5))
    assert np.all(out["status"] == 1) and np.all(out["cell"] == -1)
    assert np.allclose(out["certified"], 1.0, atol=2.5 * 2.0 ** -27 + 1e-9)
    # within the arm the same rays succeed
    ok = _lib.ray_trace(pts, tets, start, origin, direction, np.full((len(start), 1), 0.9))
    assert np.all(ok["status"] == 0)


def test_planted_gap_is_seen_by_the_checker():
    """The checker itself: removing one cube of a box opens a gap that
    segment_cover reports at the removed cube's face."""
    pts, tets = R.kuhn_box(4, 1, 1, 1.0, keep=lambda i, j, k: i != 2)
    inside, reach = R.segment_cover(pts, tets, np.array([0.0, 0.5, 0.5]), np.array([1.0, 0, 0]), 3.5)
    assert not inside and abs(reach - 2.0) < 1e-8  # intervals carry the 1e-9 barycentric tolerance
    inside, reach = R.segment_cover(pts, tets, np.array([0.0, 0.5, 0.5]), np.array([1.0, 0, 0]), 1.9)
    assert inside


def test_invalid_input_is_an_error():
    pts, tets = R.kuhn_box(1, 1, 1)
    bad = tets.copy()
    bad[0, 1] = bad[0, 0]
    with pytest.raises(RuntimeError, match="repeated vertex"):
        _lib.ray_trace(pts, bad, [0], [[0.1, 0.1, 0.1]], [[1, 0, 0]], [[0.1]])
    with pytest.raises(RuntimeError, match="out of range"):
        _lib.ray_trace(pts, tets, [99], [[0.1, 0.1, 0.1]], [[1, 0, 0]], [[0.1]])


def _turb_samples(mesh, extra, create=()):
    from nodals_amd._lib import H8

    h = H8(["--mesh", mesh, "--max-outer", "10", "--verbose", "0", *create], device=False)
    h.turb_init(extra)
    n = 12 * h.turb["wall_faces"]
    out = {
        "cell": h.get("turb_sample_cell", n * 5).astype(int).reshape(n, 5),
        "l": h.get("turb_sample_l", n * 20).reshape(n, 5, 4),
        "y": h.get("turb_sample_y", n * 5).reshape(n, 5),
        "stats": h.get("turb_sample_stats", 7),
        "heights": h.get("turb_heights", 5),
    }
    h.close()
    return out


def _positions(mesh_points, tets, s):
    return np.einsum("qmj,qmjd->qmd", s["l"], mesh_points[tets[s["cell"]]])


def test_ray_placement_equals_the_walk_on_the_helix(mesh_dir):
    """All five slots traced: every sample is the walk's point (same position
    to rounding; the cell may be the other side of a shared face)."""
    walk = _turb_samples(mesh_dir["helix"], [], ["--duct", "1"])
    ray = _turb_samples(mesh_dir["helix"], ["--sst-wall-sampling", "ray_trace", "--sst-wall-active-points", "5"],
                        ["--duct", "1"])
    assert ray["stats"][6] == 1 and walk["stats"][6] == 0 and int(ray["stats"][5]) == 0x1F
    assert np.array_equal(ray["y"], walk["y"])
    same = ray["cell"] == walk["cell"]
    assert same.mean() > 0.99
    assert np.allclose(ray["l"][same], walk["l"][same], rtol=0, atol=1e-12)
    assert ray["stats"][3] == 0


def test_one_active_slot_is_traced_and_the_others_alias_it(mesh_dir):
    ray = _turb_samples(mesh_dir["small"], ["--sst-wall-sampling", "ray_trace", "--sst-wall-active-points", "1"])
    walk = _turb_samples(mesh_dir["small"], [])
    assert int(ray["stats"][5]) == 0x04
    same = ray["cell"][:, 2] == walk["cell"][:, 2]
    assert same.mean() > 0.99
    assert np.allclose(ray["l"][same, 2], walk["l"][same, 2], rtol=0, atol=1e-12)
    for m in (0, 1, 3, 4):
        assert np.array_equal(ray["cell"][:, m], ray["cell"][:, 2])
        assert np.array_equal(ray["l"][:, m], ray["l"][:, 2])
        assert np.all(ray["y"][:, m] == ray["heights"][m])
    assert ray["stats"][0] == ray["cell"].shape[0]  # one traced sample per quadrature point


def test_local_shortening_across_the_pipe(mesh_dir):
    """Third sample 1.2 D from the wall of a D = 1 pipe: every ray leaves the
    pipe at its chord, so each is shortened to 0.90 of its certified length
    (at least 0.25 of the height) and sampled there; without local
    shortening the setup fails."""
    ys = ["--sst-profile-fit-y1", "0.1", "--sst-profile-fit-y2", "0.2", "--sst-profile-fit-y3", "1.2",
          "--sst-profile-fit-y4", "1.3", "--sst-profile-fit-y5", "1.4"]
    base = ["--sst-wall-sampling", "ray_trace", "--sst-wall-active-points", "1", *ys]
    with pytest.raises(RuntimeError, match="wall sample ray"):
        _turb_samples(mesh_dir["small"], base)
    s = _turb_samples(mesh_dir["small"], base + ["--sst-wall-ray-local-adapt", "1"])
    nq = s["cell"].shape[0]
    assert s["stats"][3] == nq
    y3 = s["y"][:, 2]
    assert np.all(y3 >= 0.25 * 1.2) and np.all(y3 < 1.2)
    assert 0.25 <= s["stats"][4] < 1.0
    # y3 / 0.9 is the certified chord: between 0.5 D (a wall point's chord
    # towards the axis is at least the radius) and D
    assert np.all(y3 / 0.9 > 0.5) and np.all(y3 / 0.9 <= 1.0 + 1e-6)
    assert np.all(s["y"][:, [0, 1, 3, 4]] == s["heights"][[0, 1, 3, 4]])


def _face_csv(path, faces, utau, header=True):
    with open(path, "w") as f:
        if header:
            f.write(f"# NODALS_FACE_UTAU_V1 iteration=10 faces={len(faces)} fit=F2_lagged\n")
        for i, ((c, lf), u) in enumerate(zip(faces, utau)):
            f.write(f"{i},{int(c)},{int(lf)},{float(u)!r},{30.0!r}\n")


def _face_context(mesh, extra):
    from nodals_amd._lib import H8

    h = H8(["--mesh", mesh, "--max-outer", "10", "--verbose", "0", "--re", "5000"], device=False)
    h.turb_init(extra)
    return h


def test_per_face_heights_from_friction_velocities(mesh_dir, tmp_path):
    """The adaptive runner's hand-over (sst_application.inc:1039-1097): the
    third sample of each wall face moves to y = 30 nu / u_tau, clamped to
    [0.02, 0.35] of the full-scale third height (given height / sample
    scale); u_tau = 0 falls back to the given height, clamped the same way.  Every placed sample carries
    its height, and the rays are traced (locally shortened if needed)."""
    from nodals_amd._lib import H8

    scale = 0.5
    base = ["--sst-wall-sampling", "ray_trace", "--sst-wall-active-points", "1", "--sst-wall-sample-scale", str(scale)]
    h0 = _face_context(mesh_dir["small"], base)
    nf = h0.turb["wall_faces"]
    faces = h0.get("turb_wall_faces", 2 * nf).reshape(nf, 2)
    y3 = h0.get("turb_heights", 5)[2]
    h0.close()
    nu = 1.0 / 5000.0
    rng = np.random.default_rng(4)
    utau = 30 * nu / (y3 / scale * rng.uniform(0.005, 0.5, nf))  # y+ 30 heights from 0.005 to 0.5 full scale
    utau[:5] = 0.0
    csv = tmp_path / "NODALS_FACE_UTAU_it10.csv"
    _face_csv(csv, faces, utau)
    h = _face_context(mesh_dir["small"], base + ["--sst-face-utau-csv", str(csv), "--sst-spalding-sample-scale",
                                                 str(scale)])
    got = h.get("turb_face_y3", nf)
    full = y3 / scale
    # the fallback height is clamped like the others
    unbounded = np.where(utau > 1e-8, 30 * nu / np.where(utau > 0, utau, 1.0), y3)
    want = np.clip(unbounded, 0.02 * full, 0.35 * full)
    assert np.allclose(got, want, rtol=1e-14, atol=0)
    st = h.get("turb_face_yplus_stats", 10)
    assert st[0] == nf and st[1] == 5
    assert st[2] == np.sum(np.abs(want - unbounded) > 1e-13) and 0 < st[2] < nf
    nq = 12 * nf
    sy = h.get("turb_sample_y", nq * 5).reshape(nf, 12, 5)
    stats = h.get("turb_sample_stats", 7)
    shortened = int(stats[3])
    # unshortened third samples sit at the face's height
    at = np.isclose(sy[:, :, 2], got[:, None], rtol=1e-15, atol=0)
    assert at.sum() == nq - shortened
    assert np.all(sy[:, :, 2] <= got[:, None] * (1 + 1e-15))
    h.close()


def test_per_face_heights_reject_mismatched_files(mesh_dir, tmp_path):
    base = ["--sst-wall-sampling", "ray_trace", "--sst-wall-active-points", "1"]
    h0 = _face_context(mesh_dir["small"], base)
    nf = h0.turb["wall_faces"]
    faces = h0.get("turb_wall_faces", 2 * nf).reshape(nf, 2)
    h0.close()
    csv = tmp_path / "f.csv"
    face = ["--sst-face-utau-csv", str(csv), "--sst-spalding-sample-scale", "0.5"]
    bad = faces.copy()
    bad[3, 0] += 1
    _face_csv(csv, bad, np.full(nf, 0.05))
    with pytest.raises(RuntimeError, match="does not match the wall faces at row 3"):
        _face_context(mesh_dir["small"], base + face)
    _face_csv(csv, faces[:-1], np.full(nf - 1, 0.05))
    with pytest.raises(RuntimeError, match="wrong number of faces"):
        _face_context(mesh_dir["small"], base + face)
    _face_csv(csv, faces, np.full(nf, 0.05))
    with pytest.raises(RuntimeError, match="need ray_trace"):
        _face_context(mesh_dir["small"], ["--sst-wall-active-points", "1"] + face)
    with pytest.raises(RuntimeError, match="omega target must equal"):
        _face_context(mesh_dir["small"], base + face + ["--sst-face-yplus-omega", "40"])
