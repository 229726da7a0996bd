"""Fiber files in the reference's text formats.

Reading follows read_input/read_fibers_file.py:19-68: '#' starts a comment,
blank lines are skipped, the first value is the number of fibers, and each
fiber is a header line "num_points [E [length [growing]]]" followed by
num_points lines "x y z".

Writing follows tstep.write_to_file (python/tstep/tstep.py:2158-2177):
  * a trajectory frame of <output>_<ID>_fibers.txt (output_txt_files=True,
    the reference default): a row of four copies of the fiber count, then per
    fiber the row (num_points, E, length, growing) and num_points rows
    (x, y, z, tension), every row written by numpy.savetxt's default '%.18e';
  * the same frame for <output>_<ID>.fibers (output_txt_files=False): the
    fiber count and the per-fiber header as plain '%s' text, the rows by
    numpy.savetxt;
  * <output>_<ID>_resume.fibers: the fiber count, then per fiber the '%s'
    header and num_points rows (x, y, z).
Output is byte-for-byte what the reference writes for the same values.

The simulation input file (options, structures) follows
read_input/read_input.py:9-191 (ReadInput).

Body vertex and link files follow read_input/read_vertex_file.py and
read_input/read_links_file.py.

Rigid bodies (clones) follow read_input/read_clones_file.py:9-48 and the
body part of tstep.write_to_file (tstep.py:2078-2123): a count, then one
"x y z s p1 p2 p3" line per body (location and orientation quaternion,
normalised on reading); trajectory frames carry a row of seven copies of the
count and, per body, a row of seven copies of its radius before its
location-orientation row; velocity frames a row of six copies of the count,
then "U Omega" per body.
"""
import ntpath

import numpy as np


def _floats(text):
    """Whitespace-separated numbers up to the first token that is not one
    (what numpy.fromstring(text, sep=' ') returns, without its warning)."""
    out = []
    for tok in text.split():
        try:
            out.append(float(tok))
        except ValueError:
            break
    return np.array(out)


def _number(options, key, default, kind):
    # an absent option and an empty value both take the default, as the
    # reference's `options.get(key) or default`
    v = options.get(key)
    return kind(v) if v else kind(default)


class ReadInput:
    """The simulation input file (read_input/read_input.py:9-191).

    One option per line, "name value...", '#' starts a comment, blank lines
    are skipped; the value is the rest of the line.  Repeated structure
    options are numbered in order (structure0, structure1, ...; likewise
    structure_fibers, structure_body_fibers, mm_surface_structure).  The
    attributes below carry the reference's names and defaults."""

    _FLOATS = [("dt", 0.0), ("eta", 1.0), ("kT", 1.0), ("solver_tolerance", 1e-08),
               ("E", 0.0), ("length_fibers", 0.0), ("epsilon", 1e-03), ("v_growth", 0.0),
               ("v_shrink", 0.0), ("rate_catastrophe", 0.0), ("rate_rescue", 0.0),
               ("rate_seed", 0.0), ("force_stall", 1.0), ("rate_catastrophe_stall", 1.0),
               ("length_min", 0.0), ("molecular_motor_radius", 1.0),
               ("molecular_motor_speed", 1.0), ("molecular_motor_force_stall", 1.0),
               ("molecular_motor_spring_constant", 1.0), ("molecular_motor_rest_length", 1.0),
               ("molecular_motor_bind_frequency", 1.0),
               ("molecular_motor_unbind_frequency", 1.0), ("molecular_motor_kernel_sigma", 1.0)]
    _INTS = [("n_steps", 0), ("initial_step", 0), ("n_save", 1), ("num_points_fibers", 0),
             ("p_fmm", 8), ("num_points_finite_diff", 0)]
    _STRINGS = [("scheme", "deterministic_forward_euler"), ("output_name", "run"),
                ("save_clones", "one_file_per_step"),
                ("mobility_vector_prod_implementation", "single_layer_double_layer")]
    _RAW = ["random_state", "seed", "force_file", "velocity_file"]
    _OR_NONE = ["force_generator_config", "molecular_motor_config",
                "molecular_motor_attached_ends", "molecular_motor_head_config",
                "molecular_motor_base_config", "molecular_motor_s_head",
                "molecular_motor_s_base"]
    _NUMBERED = ("structure", "structure_fibers", "structure_body_fibers",
                 "mm_surface_structure")

    def __init__(self, entries):
        self.entries = entries
        self.input_file = entries
        self.options = {}
        count = {k: 0 for k in self._NUMBERED}
        with open(self.input_file, "r") as f:
            for line in f:
                line = line.split("#", 1)[0].strip()
                if not line:
                    continue
                parts = line.split(None, 1)
                if len(parts) < 2:
                    # the reference's tuple unpacking fails on a bare name
                    raise ValueError("option %r has no value" % parts[0])
                option, value = parts
                if option in count:
                    n = count[option]
                    count[option] = n + 1
                    option = option + str(n)
                self.options[option] = value
        o = self.options
        for key, d in self._FLOATS:
            setattr(self, key, _number(o, key, d, float))
        for key, d in self._INTS:
            setattr(self, key, _number(o, key, d, int))
        for key, d in self._STRINGS:
            setattr(self, key, str(o.get(key) or d))
        for key in self._RAW:
            setattr(self, key, o.get(key))
        for key in self._OR_NONE:
            setattr(self, key, o.get(key) or None)
        self.periodic_length = _floats(o.get("periodic_length") or "0 0 0")
        self.plot_velocity_field = _floats(o.get("plot_velocity_field") or "None")

        def numbered(name, n, split=True):
            vals = [o.get(name + str(i)) for i in range(n)]
            return [str(v).split() if split else str(v) for v in vals]

        def ids(structs, pick):
            # file name without directory and without its 7-character suffix
            # ('.clones' / '.fibers')
            return [ntpath.split(pick(s))[1][:-7] for s in structs]

        self.structures = numbered("structure", count["structure"])
        self.structures_ID = ids(self.structures, lambda s: s[1])
        if self.initial_step > 0:
            for k, st in enumerate(self.structures):
                st[1] = "%s.%s.%s.clones" % (self.output_name, self.structures_ID[k],
                                             str(self.initial_step).zfill(8))
        self.structures_fibers = numbered("structure_fibers", count["structure_fibers"],
                                          split=False)
        self.structures_fibers_ID = ids(self.structures_fibers, lambda s: s)
        if self.initial_step > 0 and self.structures_fibers:
            # read_input.py:136-139 assigns into a string here, which raises
            raise TypeError("restarting fiber structures (initial_step > 0) is not supported "
                            "by the reference's input reader")
        self.structures_body_fibers = numbered("structure_body_fibers",
                                               count["structure_body_fibers"])
        self.structures_body_fibers_ID = ids(self.structures_body_fibers, lambda s: s[1])
        if self.initial_step > 0:
            # read_input.py:162-164: prints this and exits (status 0)
            print("ERROR, code is not ready to restart from old configurations")
            raise SystemExit()
        self.mm_surface_structure = numbered("mm_surface_structure",
                                             count["mm_surface_structure"])
        self.mm_surface_structure_ID = ids(self.mm_surface_structure, lambda s: s[0])


def read_fibers_file(name_file):
    """(info, coor): info[k] = [num_points, (E), (length), (growing)] as
    given in fiber k's header; coor = (sum num_points, 3) float64 array."""
    info, coor = [], []
    remaining = 0
    nfib = None
    with open(name_file, "r") as f:
        for line in f:
            line = line.split("#", 1)[0].strip()
            if not line:
                continue
            data = line.split()
            if nfib is None:
                nfib = int(float(data[0]))
                continue
            if remaining == 0:
                head = [int(float(data[0]))]
                if len(data) > 1:
                    head.append(float(data[1]))
                if len(data) > 2:
                    head.append(float(data[2]))
                if len(data) > 3:
                    # the reference converts a textual flag with bool(): any
                    # non-empty string, "0" included, reads as True
                    head.append(bool(data[3]))
                info.append(head)
                remaining = head[0]
            else:
                coor.append([float(data[0]), float(data[1]), float(data[2])])
                remaining -= 1
    return info, np.array(coor)


def fibers_from_file(name_file, E=None, length=None):
    """Per-fiber position arrays plus per-fiber E and length (header values
    where given, else the defaults passed here)."""
    info, coor = read_fibers_file(name_file)
    xs, Es, Ls = [], [], []
    o = 0
    for head in info:
        n = head[0]
        xs.append(coor[o:o + n].copy())
        Es.append(head[1] if len(head) > 1 else E)
        Ls.append(head[2] if len(head) > 2 else length)
        o += n
    return xs, Es, Ls


def _rows(f, a):
    np.savetxt(f, np.asarray(a, dtype=np.float64))


def write_fibers_frame(f, positions, tensions, E, lengths, growing=None, txt=True):
    """One output frame for the fibers of one type.  f: a file opened in
    binary mode; positions[k] (N_k, 3); tensions[k] (N_k,); E, lengths,
    growing: per fiber (scalars are broadcast)."""
    n = len(positions)
    E = np.broadcast_to(np.asarray(E, dtype=object), (n,))
    lengths = np.broadcast_to(np.asarray(lengths, dtype=object), (n,))
    growing = np.broadcast_to(np.asarray(0 if growing is None else growing, dtype=object), (n,))
    if txt:
        np.savetxt(f, np.ones((1, 4), dtype=int) * n)
    else:
        f.write((str(n) + "\n").encode("utf-8"))
    for k in range(n):
        x = np.asarray(positions[k], dtype=np.float64)
        if txt:
            head = np.array([x.shape[0], E[k], lengths[k], growing[k]], dtype=np.float64)
            np.savetxt(f, head[None, :])
        else:
            f.write(("%s %s %s %s \n" % (x.shape[0], E[k], lengths[k], growing[k])).encode("utf-8"))
        _rows(f, np.concatenate((x, np.asarray(tensions[k], np.float64)[:, None]), axis=1))


def write_resume_fibers(name_file, positions, E, lengths, growing=None):
    """<output>_<ID>_resume.fibers for the fibers of one type."""
    n = len(positions)
    E = np.broadcast_to(np.asarray(E, dtype=object), (n,))
    lengths = np.broadcast_to(np.asarray(lengths, dtype=object), (n,))
    growing = np.broadcast_to(np.asarray(0 if growing is None else growing, dtype=object), (n,))
    with open(name_file, "wb") as f:
        f.write(("%s \n" % n).encode("utf-8"))
        for k in range(n):
            x = np.asarray(positions[k], dtype=np.float64)
            f.write(("%s %s %s %s \n" % (x.shape[0], E[k], lengths[k], growing[k])).encode("utf-8"))
            _rows(f, x)


def _data_lines(name_file):
    """Non-empty lines of a reference text file, '#' comments removed."""
    with open(name_file, "r") as f:
        for line in f:
            line = line.split("#", 1)[0].strip()
            if line:
                yield line.split()


def read_vertex_file(name_file):
    """Body vertex file (read_input/read_vertex_file.py:7-33): a count line,
    then "x y z" per node -> (Nblobs, 3) array (the count is read, not
    enforced, as in the reference)."""
    rows = list(_data_lines(name_file))
    return np.array([[float(v) for v in r[:3]] for r in rows[1:]])


def read_links_file(name_file):
    """Links file (read_input/read_links_file.py:7-48): a count line, then
    "k k_angle x y z ax ay az" per link -> (spring constants (n, 1),
    angular spring constants (n, 1), locations (n, 3), unit axes (n, 3))."""
    rows = list(_data_lines(name_file))[1:]
    k = np.array([[float(r[0])] for r in rows])
    ka = np.array([[float(r[1])] for r in rows])
    loc = np.array([[float(v) for v in r[2:5]] for r in rows])
    ax = np.array([[float(v) for v in r[5:8]] for r in rows])
    if len(rows):
        ax = ax / np.linalg.norm(ax, axis=1, keepdims=True)
    return k, ka, loc, ax


def read_clones_file(name_file):
    """(number_of_bodies, locations (n, 3), orientations (n, 4)): the
    reference returns Quaternion objects, here their normalised entries
    (s, p1, p2, p3)."""
    n = None
    loc, ori = [], []
    with open(name_file, "r") as f:
        for line in f:
            line = line.split("#", 1)[0].strip()
            if not line:
                continue
            data = line.split()
            if n is None:
                n = int(data[0])
                continue
            q = np.array([float(v) for v in data[3:7]])
            loc.append([float(v) for v in data[0:3]])
            ori.append(q / np.linalg.norm(q))
    return n, np.array(loc), np.array(ori)


def write_clones_frame(f, locations, orientations, radius, txt=True):
    """One frame of <output>_<ID>.clones (txt) for the bodies of one type."""
    locations = np.asarray(locations, np.float64).reshape(-1, 3)
    orientations = np.asarray(orientations, np.float64).reshape(-1, 4)
    n = locations.shape[0]
    radius = np.broadcast_to(np.asarray(radius, np.float64), (n,))
    if txt:
        np.savetxt(f, np.ones((1, 7), dtype=int) * n)
    for k in range(n):
        row = np.concatenate((locations[k], orientations[k]))
        if txt:
            np.savetxt(f, np.ones((1, 7), dtype=int) * radius[k])
            np.savetxt(f, row[None, :])
        else:
            f.write(("%s %s %s %s %s %s %s \n" % tuple(row)).encode("utf-8"))


def write_body_velocities_frame(f, velocities):
    """One frame of the bodies' velocity output: U and Omega per body."""
    v = np.asarray(velocities, np.float64).reshape(-1, 6)
    np.savetxt(f, np.ones((1, 6), dtype=int) * v.shape[0])
    for k in range(v.shape[0]):
        np.savetxt(f, v[k][None, :])


def write_resume_clones(name_file, locations, orientations):
    """<output>_<ID>_resume.clones for the bodies of one type."""
    locations = np.asarray(locations, np.float64).reshape(-1, 3)
    orientations = np.asarray(orientations, np.float64).reshape(-1, 4)
    with open(name_file, "wb") as f:
        f.write(("%s \n" % locations.shape[0]).encode("utf-8"))
        for k in range(locations.shape[0]):
            row = np.concatenate((locations[k], orientations[k]))
            f.write(("%s %s %s %s %s %s %s \n" % tuple(row)).encode("utf-8"))
