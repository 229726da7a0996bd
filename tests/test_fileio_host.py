"""Reference-format fiber files: reading the reference's own example
structures (tests/golden/*.fibers, copied from python/multi_fibers/), its
input-file reader (ReadInput), and
writing trajectory frames and resume files that read back exactly."""
import io
import os

import numpy as np
import pytest

from fiber3d_amd.fileio import (fibers_from_file, read_fibers_file, write_fibers_frame,
                                write_resume_fibers)

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def test_read_reference_example_two_fibers():
    info, coor = read_fibers_file(os.path.join(GOLDEN, "two.fibers"))
    assert info == [[32, 1.0, 1.85], [32, 1.0, 1.85]]
    assert coor.shape == (64, 3)
    # both fibers lie on the z axis, the second shifted by 2 (file values)
    assert np.all(coor[:, :2] == 0.0)
    assert coor[0, 2] == 2.0 and coor[32, 2] == 4.0
    assert coor[31, 2] == 0.1499999999999999944


def test_read_skips_comments_and_blank_lines():
    info, coor = read_fibers_file(os.path.join(GOLDEN, "one.fibers"))
    assert info == [[32, 1.0, 1.0]] and coor.shape == (32, 3)
    xs, Es, Ls = fibers_from_file(os.path.join(GOLDEN, "one.fibers"))
    assert len(xs) == 1 and xs[0].shape == (32, 3) and Es == [1.0] and Ls == [1.0]


def test_growing_flag_reads_like_the_reference(tmp_path):
    """The reference converts the textual fourth header field with bool():
    '0' is a non-empty string and reads as True."""
    path = tmp_path / "growing.fibers"
    path.write_text("1\n2 0.01 1.5 0\n0 0 0\n0 0 1\n")
    info, coor = read_fibers_file(path)
    assert info == [[2, 0.01, 1.5, True]] and coor.shape == (2, 3)


def test_resume_file_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    pos = [rng.normal(size=(n, 3)) for n in (16, 24, 32)]
    name = tmp_path / "run_fibers_resume.fibers"
    write_resume_fibers(name, pos, 0.01, [1.0, 1.5, 2.0])
    info, coor = read_fibers_file(name)
    assert info == [[16, 0.01, 1.0, True], [24, 0.01, 1.5, True], [32, 0.01, 2.0, True]]
    assert np.array_equal(coor, np.concatenate(pos))   # '%.18e' round-trips FP64 exactly
    text = name.read_text().splitlines()
    assert text[0] == "3 " and text[1] == "16 0.01 1.0 0 "


def test_trajectory_frame_layouts():
    pos = [np.arange(6.0).reshape(2, 3), np.arange(9.0).reshape(3, 3) + 0.5]
    ten = [np.array([1.0, 2.0]), np.array([3.0, 4.0, 5.0])]
    buf = io.BytesIO()
    write_fibers_frame(buf, pos, ten, 0.01, [1.0, 2.0], txt=True)
    rows = np.loadtxt(io.StringIO(buf.getvalue().decode()))
    assert rows.shape == (1 + 1 + 2 + 1 + 3, 4)
    assert np.all(rows[0] == 2)
    assert np.array_equal(rows[1], [2, 0.01, 1.0, 0]) and np.array_equal(rows[4], [3, 0.01, 2.0, 0])
    assert np.array_equal(rows[2:4], np.hstack([pos[0], ten[0][:, None]]))
    assert np.array_equal(rows[5:], np.hstack([pos[1], ten[1][:, None]]))
    first = buf.getvalue().decode().splitlines()[0]
    assert first == " ".join(["2.000000000000000000e+00"] * 4)
    buf = io.BytesIO()
    write_fibers_frame(buf, pos, ten, 0.01, [1.0, 2.0], txt=False)
    lines = buf.getvalue().decode().splitlines()
    assert lines[0] == "2" and lines[1] == "2 0.01 1.0 0 "
    assert len(lines) == 1 + (1 + 2) + (1 + 3)


def test_clones_round_trip_and_frames(tmp_path):
    from fiber3d_amd.fileio import (read_clones_file, write_body_velocities_frame,
                                    write_clones_frame, write_resume_clones)
    loc = np.array([[0.5, -1.0, 2.0], [3.0, 0.25, -0.125]])
    q = np.array([[1.0, 0.0, 0.0, 0.0], [0.5, 0.5, -0.5, 0.5]])
    name = tmp_path / "run_bodies_resume.clones"
    write_resume_clones(name, loc, q)
    assert name.read_text().splitlines()[:2] == ["2 ", "0.5 -1.0 2.0 1.0 0.0 0.0 0.0 "]
    n, l2, q2 = read_clones_file(name)
    assert n == 2 and np.array_equal(l2, loc) and np.array_equal(q2, q)
    # orientations are normalised on reading, as the reference does
    (tmp_path / "u.clones").write_text("1\n0 0 0 2 0 0 0 # comment\n")
    assert np.array_equal(read_clones_file(tmp_path / "u.clones")[2], [[1.0, 0, 0, 0]])
    buf = io.BytesIO()
    write_clones_frame(buf, loc, q, 0.5, txt=True)
    rows = np.loadtxt(io.StringIO(buf.getvalue().decode()))
    assert rows.shape == (5, 7) and np.all(rows[0] == 2) and np.all(rows[1] == 0.5)
    assert np.array_equal(rows[2], np.concatenate([loc[0], q[0]]))
    buf = io.BytesIO()
    write_body_velocities_frame(buf, np.arange(12.0).reshape(2, 6))
    rows = np.loadtxt(io.StringIO(buf.getvalue().decode()))
    assert rows.shape == (3, 6) and np.all(rows[0] == 2) and np.array_equal(rows[2], np.arange(6.0, 12.0))


INPUT = """# Select scheme
scheme                                   linear_solver_fibers_and_bodies

# Time step and number of steps
dt                                       1e-02
n_steps                                  8000
n_save                                   20
initial_step                             0
eta                                      1.0     # viscosity
num_points_fibers                        32
E                                        1.0
length_fibers                            5
num_points_finite_diff                   11
rate_catastrophe_stall                   1e+03
# structure_fibers                       Structures/skipped.fibers
structure_fibers                         Structures/one.fibers
structure_fibers                         /data/run/two.fibers
structure_body_fibers                    Structures/b.vertex  Structures/b.clones Structures/b.fibers  Structures/b.links
structure                                sphere.shape Structures/sphere.clones
solver_tolerance                         1e-12
seed                                     1
output_name                              data/run1
periodic_length                          1 2.5 3
plot_velocity_field                      -10 10 50 -10 10 50 -10 10 50
"""


def test_read_input_options_and_structures(tmp_path):
    """ReadInput (read_input/read_input.py:9-191): values, the reference's
    defaults for absent options, numbered structures and their IDs."""
    from fiber3d_amd.fileio import ReadInput
    f = tmp_path / "data.main"
    f.write_text(INPUT)
    r = ReadInput(str(f))
    assert r.scheme == "linear_solver_fibers_and_bodies"
    assert (r.dt, r.n_steps, r.n_save, r.initial_step) == (1e-2, 8000, 20, 0)
    assert (r.eta, r.E, r.length_fibers, r.num_points_fibers) == (1.0, 1.0, 5.0, 32)
    assert r.num_points_finite_diff == 11 and r.rate_catastrophe_stall == 1e3
    assert r.solver_tolerance == 1e-12 and r.seed == "1" and r.random_state is None
    assert r.output_name == "data/run1"
    # defaults of options the file does not set
    assert (r.kT, r.epsilon, r.force_stall, r.p_fmm) == (1.0, 1e-3, 1.0, 8)
    assert r.save_clones == "one_file_per_step" and r.molecular_motor_config is None
    assert r.mobility_vector_prod_implementation == "single_layer_double_layer"
    np.testing.assert_array_equal(r.periodic_length, [1.0, 2.5, 3.0])
    np.testing.assert_array_equal(r.plot_velocity_field, [-10, 10, 50] * 3)
    assert r.structures_fibers == ["Structures/one.fibers", "/data/run/two.fibers"]
    assert r.structures_fibers_ID == ["one", "two"]
    assert r.structures_body_fibers == [["Structures/b.vertex", "Structures/b.clones",
                                         "Structures/b.fibers", "Structures/b.links"]]
    assert r.structures_body_fibers_ID == ["b"]
    assert r.structures == [["sphere.shape", "Structures/sphere.clones"]]
    assert r.structures_ID == ["sphere"]
    assert r.mm_surface_structure == [] and r.options["structure_fibers1"] == "/data/run/two.fibers"


def test_read_input_defaults_and_restart(tmp_path):
    from fiber3d_amd.fileio import ReadInput
    f = tmp_path / "empty.main"
    f.write_text("# nothing\n\n")
    r = ReadInput(str(f))
    assert (r.dt, r.eta, r.n_save, r.solver_tolerance) == (0.0, 1.0, 1, 1e-8)
    assert r.scheme == "deterministic_forward_euler" and r.output_name == "run"
    np.testing.assert_array_equal(r.periodic_length, [0.0, 0.0, 0.0])
    assert r.plot_velocity_field.size == 0
    # restarting: bodies read their clones from the output directory; the
    # reference stops there (read_input.py:113-116, :162-164)
    f.write_text("initial_step 40\noutput_name out/run\nstructure s.vertex in/s.clones\n")
    with pytest.raises(SystemExit):
        ReadInput(str(f))
    f.write_text("initial_step 40\nstructure_fibers in/f.fibers\n")
    with pytest.raises(TypeError):
        ReadInput(str(f))
    f.write_text("dt\n")
    with pytest.raises(ValueError):
        ReadInput(str(f))


def test_read_input_reference_example():
    """The reference's own example input (python/multi_fibers/one_body/
    data.main, copied as a fixture)."""
    from fiber3d_amd.fileio import ReadInput
    r = ReadInput(os.path.join(GOLDEN, "one_body_data.main"))
    assert r.scheme == "linear_solver_fibers_and_bodies"
    assert (r.dt, r.n_steps, r.n_save, r.eta, r.E) == (1e-2, 8000, 20, 1.0, 1.0)
    assert (r.num_points_fibers, r.length_fibers, r.num_points_finite_diff) == (32, 5.0, 11)
    assert (r.v_shrink, r.rate_catastrophe, r.rate_rescue) == (0.065, 1e-5, 0.014)
    assert r.solver_tolerance == 1e-12 and r.output_name == "data/run1"
    assert r.structures_fibers == []         # commented out in the file
    assert r.structures_body_fibers == [["Structures/body_N_1.vertex",
                                         "Structures/body_N_1.clones",
                                         "Structures/body_N_1.fibers",
                                         "Structures/body_N_1.links"]]
    assert r.structures_body_fibers_ID == ["body_N_1"]


def test_vertex_and_links_files(tmp_path):
    """Body vertex and link files in the reference's formats
    (read_input/read_vertex_file.py, read_links_file.py): comments and blank
    lines skipped, link axes normalised."""
    from fiber3d_amd.fileio import read_links_file, read_vertex_file
    v = tmp_path / "body.vertex"
    v.write_text("# nodes\n3\n0 0 1\n\n1.5 -2 0.25  # a comment\n0 1e-3 -1\n")
    np.testing.assert_array_equal(read_vertex_file(str(v)),
                                  [[0, 0, 1], [1.5, -2, 0.25], [0, 1e-3, -1]])
    l = tmp_path / "body.links"
    l.write_text("2\n10 5 0 0 0.5 0 0 2\n# second\n1 0.5 0.3 0 0.4 3 4 0\n")
    k, ka, loc, ax = read_links_file(str(l))
    np.testing.assert_array_equal(k, [[10.0], [1.0]])
    np.testing.assert_array_equal(ka, [[5.0], [0.5]])
    np.testing.assert_array_equal(loc, [[0, 0, 0.5], [0.3, 0, 0.4]])
    np.testing.assert_allclose(ax, [[0, 0, 1], [0.6, 0.8, 0]], rtol=0, atol=1e-15)


def test_stepper_from_input_rejects_unset_stencil_and_missing_length(tmp_path):
    """stepper_from_input needs num_points_finite_diff (the reference would
    build one-point stencils from an unset 0) and a positive length for every
    fiber; both are checked before any device work."""
    from fiber3d_amd.fileio import ReadInput
    from fiber3d_amd.tstep import stepper_from_input
    f = tmp_path / "data.main"
    f.write_text("dt 1e-3\neta 1.0\nE 0.5\nstructure_fibers %s\n"
                 % os.path.join(GOLDEN, "two.fibers"))
    with pytest.raises(ValueError, match="num_points_finite_diff"):
        stepper_from_input(ReadInput(str(f)), None)
    fib = tmp_path / "nolen.fibers"
    fib.write_text("1\n5 0.5\n" + "".join(f"{0.1 * k} 0 0\n" for k in range(5)))
    f.write_text("dt 1e-3\nnum_points_finite_diff 1\nstructure_fibers %s\n" % fib)
    with pytest.raises(ValueError, match="positive length"):
        stepper_from_input(ReadInput(str(f)), None)
