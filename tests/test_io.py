"""Problems on disk (include/vba_io.h), on the CPU.

COLMAP cases restate the reference's own tests of its reader and writer
(crates/io/tests/colmap.rs) on the fixture they read
(crates/io/tests/fixtures/colmap_text, copied under tests/golden/colmap_text)
and on the binary records their helpers build (:436-628). The problem dump
is checked against the format of pipelines/slam/src/ba_problem_io.rs and
for exact reloads; the local window against ba_gpu_replay.rs:43-66."""
import os
import struct

import numpy as np
import pytest

import vba
from vba import abi

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "colmap_text")


def read_text(tmp_path, cameras, images, points):
    d = tmp_path / "model"
    d.mkdir(exist_ok=True)
    (d / "cameras.txt").write_text(cameras)
    (d / "images.txt").write_text(images)
    (d / "points3D.txt").write_text(points)
    return vba.Model.read_colmap(d)


def write_bin(tmp_path, cameras, images, points, name="bin"):
    d = tmp_path / name
    d.mkdir(exist_ok=True)
    (d / "cameras.bin").write_bytes(cameras)
    (d / "images.bin").write_bytes(images)
    (d / "points3D.bin").write_bytes(points)
    return d


def unit(q):
    q = np.asarray(q, np.float64)
    n = np.sqrt(((q[1] * q[1] + q[2] * q[2]) + q[3] * q[3]) + q[0] * q[0])
    return q / n


# ---- the reference's binary fixture builders (colmap.rs:436-628) ----------

def u32(v): return struct.pack("<I", v)
def i32(v): return struct.pack("<i", v)
def u64(v): return struct.pack("<Q", v)
def i64(v): return struct.pack("<q", v)
def f64(v): return struct.pack("<d", v)


def camera_bin():
    b = u64(2)
    b += u32(1) + i32(1) + u64(640) + u64(480) + b"".join(f64(v) for v in (500.0, 510.0, 320.0, 240.0))
    b += u32(2) + i32(0) + u64(800) + u64(600) + b"".join(f64(v) for v in (700.0, 400.0, 300.0))
    return b


def image_header(image_id, camera_id, name):
    return u32(image_id) + b"".join(f64(v) for v in (1.0, 0.0, 0.0, 0.0, 0.1, 0.2, 0.3)) + u32(camera_id) + \
        name.encode() + b"\0"


def images_bin():
    b = u64(2)
    b += image_header(10, 1, "image_a.jpg") + u64(3)
    b += f64(320.0) + f64(240.0) + i64(1000) + f64(10.0) + f64(20.0) + i64(-1) + f64(400.0) + f64(200.0) + i64(1001)
    b += image_header(11, 2, "image_b.jpg") + u64(1) + f64(123.0) + f64(456.0) + i64(1001)
    return b


def point3d(pid, xyz, track):
    b = u64(pid) + b"".join(f64(v) for v in xyz) + bytes([255, 128, 0]) + f64(0.25) + u64(len(track))
    return b + b"".join(u32(i) + u32(k) for i, k in track)


def points3d_bin():
    return u64(2) + point3d(1000, (1.0, 2.0, 3.0), [(10, 0)]) + point3d(1001, (-1.0, 0.5, 4.0), [(10, 2), (11, 0)])


def points3d_bin_with_track_length(n):
    return u64(1) + u64(1000) + b"".join(f64(v) for v in (1.0, 2.0, 3.0)) + bytes([255, 128, 0]) + f64(0.25) + u64(n)


# writable_map (colmap.rs:459-538) as COLMAP text
WRITABLE = (
    "1 PINHOLE 640 480 500 510 320 240\n2 SIMPLE_PINHOLE 800 600 700 400 300\n",
    "10 1 0 0 0 0.1 0.2 0.3 1 a.jpg\n320 240 1000 10 20 -1 400 200 1001\n"
    "11 1 0 0 0 0.1 0.2 0.3 2 b.jpg\n123 456 1001\n",
    "1000 1 2 3 0 0 0 0\n1001 -1 0.5 4 0 0 0 0\n",
)


# ---- readers ------------------------------------------------------------------

def test_reads_reference_text_fixture():
    """parses_colmap_{cameras,points3d,images}_txt and
    reads_colmap_text_model_from_directory (colmap.rs:42-52, 90-99, 143-156,
    226-236)."""
    p, w = vba.Model.read_colmap(GOLDEN).problem()
    assert w is None
    # two cameras in use: one rig sensor each, ascending camera id
    assert [s.camera.model for s in p.sensors] == [abi.CAMERA_PINHOLE, abi.CAMERA_SIMPLE_PINHOLE]
    assert list(p.sensors[0].camera.params[:4]) == [500.0, 510.0, 320.0, 240.0]
    assert list(p.sensors[1].camera.params[:3]) == [700.0, 400.0, 300.0]
    assert (p.sensors[0].camera.width, p.sensors[0].camera.height) == (640, 480)
    assert list(p.sensors[0].sensor_from_rig) == [1, 0, 0, 0, 0, 0, 0]
    assert list(p.landmark_ids) == [1000, 1001]
    assert p.landmarks[0, 0] == 1.0 and p.landmarks[1, 2] == 4.0
    assert list(p.pose_ids) == [10, 11]
    np.testing.assert_array_equal(p.poses[0], [1, 0, 0, 0, 0.1, 0.2, 0.3])
    np.testing.assert_array_equal(p.poses[1, :4], unit([0.9238795325, 0, 0.3826834324, 0]))
    np.testing.assert_array_equal(p.poses[1, 4:], [1.0, 2.0, 3.0])
    # image 10 has three keypoints, two observed (keypoints 0 and 2)
    assert len(p.obs_pose_ids) == 0
    assert list(p.rig_pose_ids) == [10, 10, 11]
    assert list(p.rig_landmark_ids) == [1000, 1001, 1001]
    np.testing.assert_array_equal(p.rig_xy, [[100, 120], [300, 320], [10, 20]])
    assert list(p.rig_sensor) == [0, 0, 1]
    # the SfM's gauge: the first image and the one farthest from it
    assert list(p.pose_fixed) == [1, 1]


def test_reads_binary_records(tmp_path):
    """parses_colmap_*_bin, reads_colmap_binary_model_from_directory
    (colmap.rs:54-65, 101-110, 158-171, 238-256)."""
    d = write_bin(tmp_path, camera_bin(), images_bin(), points3d_bin())
    p, _ = vba.Model.read_colmap(d, binary=True).problem()
    assert [s.camera.model for s in p.sensors] == [abi.CAMERA_PINHOLE, abi.CAMERA_SIMPLE_PINHOLE]
    assert list(p.sensors[1].camera.params[:3]) == [700.0, 400.0, 300.0]
    assert list(p.pose_ids) == [10, 11] and list(p.landmark_ids) == [1000, 1001]
    assert p.landmarks[1, 2] == 4.0
    assert list(p.rig_landmark_ids) == [1000, 1001, 1001]
    np.testing.assert_array_equal(p.rig_xy, [[320, 240], [400, 200], [123, 456]])


def expect_format_error(fn, fragment):
    with pytest.raises(vba.IoError) as e:
        fn()
    assert e.value.code == abi.ERR_FORMAT, e.value
    assert fragment in e.value.message, e.value.message
    return e.value


def test_rejects_huge_binary_counts_before_allocation(tmp_path):
    """rejects_huge_colmap_binary_counts_before_allocation (colmap.rs:173-199)."""
    good_c, good_i, good_p = camera_bin(), images_bin(), points3d_bin()
    cases = [
        ((u64(2**64 - 1), good_i, good_p), "cameras.bin", "camera_count"),
        ((good_c, u64(2**64 - 1), good_p), "images.bin", "image_count"),
        ((good_c, u64(1) + image_header(10, 1, "") + u64(2**64 - 1), good_p), "images.bin", "point_count"),
        ((good_c, good_i, u64(2**64 - 1)), "points3D.bin", "point3D_count"),
        ((good_c, good_i, points3d_bin_with_track_length(2**64 - 1)), "points3D.bin", "track_length"),
    ]
    for k, (files, name, field) in enumerate(cases):
        d = write_bin(tmp_path, *files, name=f"c{k}")
        err = expect_format_error(lambda: vba.Model.read_colmap(d, binary=True), field)
        assert name in err.message


def test_rejects_truncated_binary_counts(tmp_path):
    """rejects_truncated_colmap_binary_counts_from_remaining_bytes
    (colmap.rs:201-224)."""
    good_c, good_i, good_p = camera_bin(), images_bin(), points3d_bin()
    cases = [
        ((u64(1), good_i, good_p), "cameras.bin"),
        ((good_c, u64(1), good_p), "images.bin"),
        ((good_c, u64(1) + image_header(10, 1, "") + u64(1), good_p), "images.bin"),
        ((good_c, good_i, u64(1)), "points3D.bin"),
        ((good_c, good_i, points3d_bin_with_track_length(1)), "points3D.bin"),
    ]
    for k, (files, name) in enumerate(cases):
        d = write_bin(tmp_path, *files, name=f"t{k}")
        err = expect_format_error(lambda: vba.Model.read_colmap(d, binary=True), "remaining")
        assert name in err.message


def test_lone_equirectangular_camera(tmp_path):
    """parses_a_lone_equirectangular_camera_bin (colmap.rs:67-88): the model
    reads, its camera is not one the solver takes, and the count is checked
    against what is left of the file."""
    cam = u64(1) + u32(1) + i32(17) + u64(4096) + u64(2048) + f64(4096.0) + f64(2048.0)
    m = vba.Model.read_colmap(write_bin(tmp_path, cam, u64(0), u64(0), name="eq"), binary=True)
    with pytest.raises(vba.IoError) as e:
        m.problem()
    assert e.value.code == abi.ERR_UNSUPPORTED_CAMERA_MODEL
    d = write_bin(tmp_path, cam[:-8], u64(0), u64(0), name="eq2")
    expect_format_error(lambda: vba.Model.read_colmap(d, binary=True), "remaining")


def test_trailing_bytes_and_bad_names_are_rejected(tmp_path):
    d = write_bin(tmp_path, camera_bin() + b"\0", images_bin(), points3d_bin(), name="tr")
    expect_format_error(lambda: vba.Model.read_colmap(d, binary=True), "trailing")
    bad = images_bin().replace(b"image_a.jpg", b"image_\xff.jpg")
    d = write_bin(tmp_path, camera_bin(), bad, points3d_bin(), name="utf")
    expect_format_error(lambda: vba.Model.read_colmap(d, binary=True), "UTF-8")
    bad = images_bin().replace(b"image_b.jpg\0", b"image_b.jpg")
    d = write_bin(tmp_path, camera_bin(), bad, points3d_bin(), name="unt")
    expect_format_error(lambda: vba.Model.read_colmap(d, binary=True), "images.bin")


def test_text_line_rules(tmp_path):
    """Rust's number grammar and the readers' line rules: comments and blank
    lines skipped, CRLF accepted, short lines and stray tokens rejected."""
    m = read_text(tmp_path, "# c\r\n\r\n1 PINHOLE 640 480 +500 5.1e2 320. .24e3\r\n",
                  "# i\n10 1 0 0 0 +0.1 inf -0 1 n.jpg\n1 2 7\n", "7 1 2 3 x y z w\n")
    p, _ = m.problem()
    np.testing.assert_array_equal(p.camera.params[:4], [500.0, 510.0, 320.0, 240.0])
    assert np.isinf(p.poses[0, 5]) and p.poses[0, 6] == 0.0 and np.signbit(p.poses[0, 6])
    assert list(p.obs_landmark_ids) == [7]
    for cams, imgs, pts, frag in [
        ("1 PINHOLE 640 480\n", "", "", "cameras.txt"),
        ("1 PINHOLE 640 480 0x10 1 1 1\n", "", "", "float"),
        ("1 PINHOLE 640 480 1e 1 1 1\n", "", "", "float"),
        ("1 PINHOLE 640 -480 1 1 1 1\n", "", "", "integer"),
        ("", "10 1 0 0 0 0 0 0 1\n", "", "images.txt"),
        ("", "10 1 0 0 0 0 0 0 1 n.jpg\n1 2\n", "", "images.txt"),
        ("", "", "7 1 2 3 0 0 0\n", "points3D.txt"),
    ]:
        expect_format_error(lambda: read_text(tmp_path, cams, imgs, pts), frag)


def test_missing_files_are_io_errors(tmp_path):
    with pytest.raises(vba.IoError) as e:
        vba.Model.read_colmap(tmp_path / "absent")
    assert e.value.code == abi.ERR_IO
    with pytest.raises(vba.IoError) as e:
        vba.Model.read_ba_problem(tmp_path / "absent.txt")
    assert e.value.code == abi.ERR_IO


# ---- writers --------------------------------------------------------------------

def test_formats_text_sections(tmp_path):
    """formats_colmap_text_model_sections (colmap.rs:340-355)."""
    m = read_text(tmp_path, *WRITABLE)
    out = tmp_path / "out"
    m.write_colmap(out)
    cameras = (out / "cameras.txt").read_text()
    images = (out / "images.txt").read_text()
    points = (out / "points3D.txt").read_text()
    assert "1 PINHOLE 640 480 500 510 320 240" in cameras
    assert "10 1 0 0 0 0.1 0.2 0.3 1 image_10.jpg" in images
    assert "320 240 1000 10 20 -1 400 200 1001" in images
    assert "1000 1 2 3 255 255 255 0 10 0" in points
    assert "1001 -1 0.5 4 255 255 255 0 10 2 11 0" in points
    assert cameras.startswith("# CAMERA_ID MODEL WIDTH HEIGHT PARAMS[]\n")
    assert images.startswith("# IMAGE_ID QW QX QY QZ TX TY TZ CAMERA_ID NAME\n# POINTS2D[] as X Y POINT3D_ID\n")
    assert points.startswith("# POINT3D_ID X Y Z R G B ERROR TRACK[] as IMAGE_ID POINT2D_IDX\n")


def test_text_and_binary_round_trips(tmp_path):
    """writes_and_reads_colmap_{text,binary}_model_round_trip
    (colmap.rs:357-418): the binary writer mirrors the text writer record
    for record, and points3D.bin carries the same tracks."""
    m = read_text(tmp_path, *WRITABLE)
    m.write_colmap(tmp_path / "text")
    m.write_colmap(tmp_path / "binary", binary=True)
    a, _ = vba.Model.read_colmap(tmp_path / "text").problem()
    b, _ = vba.Model.read_colmap(tmp_path / "binary", binary=True).problem()
    src, _ = m.problem()
    for x in (a, b):
        for f in ("pose_ids", "poses", "pose_fixed", "landmark_ids", "landmarks", "rig_pose_ids",
                  "rig_landmark_ids", "rig_xy", "rig_sensor"):
            np.testing.assert_array_equal(getattr(x, f), getattr(src, f), err_msg=f)
    assert list(b.rig_landmark_ids) == [1000, 1001, 1001]
    size = os.path.getsize(tmp_path / "binary" / "points3D.bin")
    assert size == 8 + 2 * 51 + 3 * 8
    imgs = (tmp_path / "binary" / "images.bin").read_bytes()
    assert b"image_10.jpg\0" in imgs and b"image_11.jpg\0" in imgs


def test_binary_writer_rejects_ids_beyond_u32(tmp_path):
    """binary_writer_rejects_ids_beyond_colmap_u32_range (colmap.rs:420-434)."""
    big = 2**32
    cams = f"1 PINHOLE 640 480 500 510 320 240\n{big} SIMPLE_PINHOLE 800 600 700 400 300\n"
    imgs = WRITABLE[1].replace("0.3 2 b.jpg", f"0.3 {big} b.jpg")
    m = read_text(tmp_path, cams, imgs, WRITABLE[2])
    with pytest.raises(vba.IoError) as e:
        m.write_colmap(tmp_path / "never", binary=True)
    assert e.value.code == abi.ERR_EXPORT
    assert not (tmp_path / "never").exists()


def test_camera_records_are_rewritten_as_the_reference_does(tmp_path):
    """colmap_camera_record (colmap/mod.rs:1297-1311): a PINHOLE with an
    all-zero distortion tail is a PINHOLE; one with a radial tail is an
    OPENCV with p1 = p2 = 0; a model with the wrong parameter count cannot
    be written."""
    p = vba.BundleProblem(
        camera=vba.camera(abi.CAMERA_PINHOLE, [500, 510, 320, 240, 0.0, 0.0], 640, 480),
        pose_ids=[1], poses=[[1, 0, 0, 0, 0, 0, 0]], pose_fixed=[1], landmark_ids=[5],
        landmarks=[[0, 0, 4]], landmark_fixed=[0], obs_pose_ids=[1], obs_landmark_ids=[5], obs_xy=[[320, 240]])
    vba.Model.from_problem(p).write_colmap(tmp_path / "a")
    assert "1 PINHOLE 640 480 500 510 320 240\n" in (tmp_path / "a" / "cameras.txt").read_text()
    p.camera = vba.camera(abi.CAMERA_PINHOLE, [500, 510, 320, 240, 0.01, -0.002], 640, 480)
    vba.Model.from_problem(p).write_colmap(tmp_path / "b")
    assert "1 OPENCV 640 480 500 510 320 240 0.01 -0.002 0 0\n" in (tmp_path / "b" / "cameras.txt").read_text()
    p.camera = vba.camera(abi.CAMERA_SIMPLE_RADIAL, [500, 320, 240], 640, 480)
    with pytest.raises(vba.IoError) as e:
        vba.Model.from_problem(p).write_colmap(tmp_path / "c")
    assert e.value.code == abi.ERR_EXPORT


def test_gauge_fixes_first_and_farthest_image(tmp_path):
    """refinement.rs:316-352: the first registered image and the one whose
    camera centre is farthest from it (the first such on a tie)."""
    imgs = "".join(f"{i} 1 0 0 0 {-x} 0 0 1 n.jpg\n1 1 7\n" for i, x in [(3, 0.0), (5, 2.0), (8, -4.0), (9, 4.0)])
    p, _ = read_text(tmp_path, "1 PINHOLE 640 480 500 500 320 240\n", imgs, "7 0 0 5 0 0 0 0\n").problem()
    assert list(p.pose_ids) == [3, 5, 8, 9]
    assert list(p.pose_fixed) == [1, 0, 1, 0]
    # one camera in use: monocular observations
    assert len(p.obs_pose_ids) == 4 and len(p.rig_pose_ids) == 0


# ---- the problem dump (ba_problem_io.rs) ------------------------------------------

DUMP = """camera 640 480 500 510.5 320 240
poses 3
7 2 0 0 0 0.5 0 0 1
4 1 0 0 0 0 0 0 0
7 1 0 0 0 0.25 0 0 0
landmarks 2
11 0 0 5
10 1 -2 6
observations 3 1
4 10 300.5 200 0.5
7 11 320 240 1
4 11 310 250 2
"""


def test_reads_problem_dump(tmp_path):
    f = tmp_path / "p.txt"
    f.write_text(DUMP)
    p, w = vba.Model.read_ba_problem(f).problem()
    assert p.camera.model == abi.CAMERA_PINHOLE and p.camera.num_params == 4
    assert (p.camera.width, p.camera.height) == (640, 480)
    assert list(p.camera.params[:4]) == [500.0, 510.5, 320.0, 240.0]
    # BTreeMap order; a repeated id keeps its last pose, fix_pose only adds
    assert list(p.pose_ids) == [4, 7]
    np.testing.assert_array_equal(p.poses[1], [1, 0, 0, 0, 0.25, 0, 0])
    assert list(p.pose_fixed) == [0, 1]
    assert list(p.landmark_ids) == [10, 11]
    assert list(p.obs_pose_ids) == [4, 7, 4] and list(p.obs_landmark_ids) == [10, 11, 11]
    np.testing.assert_array_equal(w, [0.5, 1.0, 2.0])
    assert p.declined_features & abi.FEATURE_OBSERVATION_WEIGHTS


def test_problem_dump_errors(tmp_path):
    f = tmp_path / "p.txt"
    for text in ["", "camera 1 2 3 4 5\n", "kamera 1 2 3 4 5 6\n", DUMP.replace("poses 3", "poses x"),
                 DUMP.replace("11 0 0 5", "11 0 0"), DUMP.replace("observations 3 1", "observations 4 1")]:
        f.write_text(text)
        with pytest.raises(vba.IoError) as e:
            vba.Model.read_ba_problem(f)
        assert e.value.code == abi.ERR_FORMAT, text


def test_problem_dump_reloads_exactly(tmp_path):
    """Floats are written in the shortest text that reads back to the same
    double (ba_problem_io.rs:5-6), quaternions are normalised on reading as
    UnitQuaternion::from_quaternion does."""
    s = vba.synthesize(abi.SYNTH_CIRCLE, num_cameras=6, num_landmarks=40, seed=5)
    s.problem.camera = vba.camera(abi.CAMERA_PINHOLE, [500.25, 499.75, 320.5, 240.125], 640, 480)
    rng = np.random.default_rng(0)
    weights = rng.random(len(s.problem.obs_pose_ids)) * 1e-3
    m = vba.Model.from_problem(s.problem, weights)
    m.write_ba_problem(tmp_path / "p.txt")
    text = (tmp_path / "p.txt").read_text()
    assert "e" not in text.replace("camera", "").replace("poses", "").replace("landmarks", "").replace(
        "observations", "")
    p, w = vba.Model.read_ba_problem(tmp_path / "p.txt").problem()
    np.testing.assert_array_equal(p.pose_ids, s.problem.pose_ids)
    np.testing.assert_array_equal(p.poses[:, 4:], s.problem.poses[:, 4:])
    for q_in, q_out in zip(s.problem.poses[:, :4], p.poses[:, :4]):
        np.testing.assert_array_equal(q_out, unit(q_in))
    np.testing.assert_array_equal(p.pose_fixed, s.problem.pose_fixed)
    np.testing.assert_array_equal(p.landmarks, s.problem.landmarks)
    np.testing.assert_array_equal(p.obs_xy, s.problem.obs_xy)
    np.testing.assert_array_equal(w, weights)
    # a second write differs from the first only where the reload
    # normalised a quaternion again
    vba.Model.read_ba_problem(tmp_path / "p.txt").write_ba_problem(tmp_path / "q.txt")
    first, second = text.splitlines(), (tmp_path / "q.txt").read_text().splitlines()
    assert len(first) == len(second)
    np_ = len(p.pose_ids)
    for k, (a, b) in enumerate(zip(first, second)):
        if 2 <= k < 2 + np_:
            a, b = a.split(), b.split()
            assert a[:1] + a[5:] == b[:1] + b[5:]
            np.testing.assert_array_equal([float(x) for x in b[1:5]], unit([float(x) for x in a[1:5]]))
        else:
            assert a == b


def test_shortest_float_text(tmp_path):
    f = tmp_path / "p.txt"
    vals = [0.1, 1.0, -0.0, 1e-7, 123456.789, 1e21, 0.30000000000000004, 5e-324]
    f.write_text("camera 1 1 1 1 1 1\nposes 0\nlandmarks %d\n" % len(vals) +
                 "".join(f"{i} {v!r} 0 0\n" for i, v in enumerate(vals)) + "observations 0 0\n")
    vba.Model.read_ba_problem(f).write_ba_problem(tmp_path / "o.txt")
    lines = (tmp_path / "o.txt").read_text().splitlines()[3:3 + len(vals)]
    got = [ln.split()[1] for ln in lines]
    assert got[:7] == ["0.1", "1", "-0", "0.0000001", "123456.789", "1000000000000000000000", "0.30000000000000004"]
    assert float(got[7]) == 5e-324 and "e" not in got[7]


def test_local_window(tmp_path):
    """ba_gpu_replay --local 2: the last two poses stay variable, the
    landmarks they see and all observations of those are kept, then the poses
    those observations use, the others fixed."""
    f = tmp_path / "p.txt"
    f.write_text("camera 640 480 500 500 320 240\nposes 4\n" +
                 "".join(f"{i} 1 0 0 0 {i} 0 0 {int(i == 1)}\n" for i in (1, 2, 3, 4)) +
                 "landmarks 4\n" + "".join(f"{l} 0 0 5\n" for l in (10, 11, 12, 13)) +
                 "observations 6 0\n1 10 1 1\n2 11 1 1\n3 11 1 1\n4 12 1 1\n2 10 1 1\n1 13 1 1\n")
    m = vba.Model.read_ba_problem(f)
    m.keep_local_window(2)
    p, _ = m.problem()
    assert list(p.landmark_ids) == [11, 12]
    assert list(p.obs_pose_ids) == [2, 3, 4] and list(p.obs_landmark_ids) == [11, 11, 12]
    assert list(p.pose_ids) == [2, 3, 4]
    assert list(p.pose_fixed) == [1, 0, 0]


def test_set_state_and_fixed(tmp_path):
    m = read_text(tmp_path, *WRITABLE)
    p, _ = m.problem()
    poses = p.poses.copy()
    poses[:, 4] += 1.5
    lms = p.landmarks * 2.0
    m.set_state(poses, lms)
    m.set_fixed(np.array([0, 1], np.uint8))
    q, _ = m.problem()
    np.testing.assert_array_equal(q.poses, poses)
    np.testing.assert_array_equal(q.landmarks, lms)
    assert list(q.pose_fixed) == [0, 1]
    with pytest.raises(ValueError):
        m.set_state(poses[:1])
