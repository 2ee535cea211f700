"""GPUMD input files for a SUS2-SH run: the `potential` line of run.in and the extended-XYZ
model file, read the way GPUMD reads them for this backend.

  run.in   `potential <file.mtp> <species_count element symbols> [sus2_* options]`: the
           symbols give the species index of each element (its position on the line) and the
           tokens after them are the backend options (src/model/read_xyz.cu:452-495,
           src/force/force.cu:112-114).
  model.xyz  line 1: atom count; line 2: `pbc="T T T"`, `Lattice="ax ay az bx by bz cx cy
           cz"` (cell vectors a, b, c), `Properties=species:S:1:pos:R:3[:...]`; then one row
           per atom (src/model/read_xyz.cu:154-330).  Keys are case-insensitive; pbc defaults
           to periodic.
"""
import shlex

import numpy as np


def species_count_of(mtp_path):
    """species_count of a SUS2 MTP file (first `species_count =` line)."""
    with open(mtp_path) as fh:
        first = fh.readline().split()
        if not first or first[0] != "MTP":
            raise ValueError(f"{mtp_path}: not an MTP potential file")
        for line in fh:
            tok = line.split()
            if len(tok) >= 3 and tok[0] == "species_count" and tok[1] == "=":
                return int(tok[2])
    raise ValueError(f"{mtp_path}: SUS2 MTP potential file is missing a valid species_count.")


def potential_line(run_in_path, mtp_path=None):
    """(model file, element symbols, backend options) of the first `potential` line of
    run.in.  `mtp_path` replaces the file named on the line (its species_count is read from
    whichever file is used)."""
    with open(run_in_path) as fh:
        for line in fh:
            tok = line.split("#", 1)[0].split()
            if len(tok) >= 2 and tok[0] == "potential":
                path = mtp_path or tok[1]
                S = species_count_of(path)
                if len(tok) < 2 + S:
                    raise ValueError("SUS2 MTP potential line has fewer element symbols than "
                                     "species_count.")
                return path, tok[2:2 + S], tok[2 + S:]
    raise ValueError("There is no 'potential' keyword in run.in.")


def _comment_fields(line):
    out = {}
    for item in shlex.split(line):
        if "=" in item:
            k, v = item.split("=", 1)
            out[k.lower()] = v
    return out


def read_xyz(path):
    """Extended XYZ model file -> dict(cell [3, 3] rows = cell vectors a, b, c, pbc (3 ints),
    symbols [N], positions [3, N] and, when present, masses [N] and velocities [3, N])."""
    with open(path) as fh:
        first = fh.readline().split()
        if len(first) != 1:
            raise ValueError("The first line for the xyz file should have one value.")
        n = int(first[0])
        if n < 2:
            raise ValueError("Number of atoms should >= 2.")
        fields = _comment_fields(fh.readline())
        rows = [fh.readline().split() for _ in range(n)]
    pbc = [1, 1, 1]
    if "pbc" in fields:
        flags = fields["pbc"].lower().split()
        if len(flags) != 3 or any(f not in ("t", "f") for f in flags):
            raise ValueError("pbc should hold three T/F flags.")
        pbc = [1 if f == "t" else 0 for f in flags]
    if "lattice" not in fields:
        raise ValueError("'lattice' is missing in the model file.")
    lat = [float(v) for v in fields["lattice"].split()]
    if len(lat) != 9:
        raise ValueError("Lattice should hold nine numbers.")
    cell = np.array(lat).reshape(3, 3)
    if "properties" not in fields:
        raise ValueError("'species' or 'properties' is missing in the model file.")
    spec = fields["properties"].split(":")
    cols, off = {}, 0
    for k in range(0, len(spec) - 2, 3):
        name, count = spec[k].lower(), int(spec[k + 2])
        cols[name] = (off, count)
        off += count
    if "species" not in cols or "pos" not in cols:
        raise ValueError("'species' or 'properties' is missing in the model file.")
    if any(len(r) < off for r in rows):
        raise ValueError("an atom row has fewer columns than Properties lists")
    s0 = cols["species"][0]
    p0 = cols["pos"][0]
    out = dict(cell=cell, pbc=tuple(pbc), symbols=[r[s0] for r in rows],
               positions=np.array([[float(r[p0 + d]) for r in rows] for d in range(3)]))
    if "mass" in cols:
        out["masses"] = np.array([float(r[cols["mass"][0]]) for r in rows])
    if "vel" in cols:
        v0 = cols["vel"][0]
        out["velocities"] = np.array([[float(r[v0 + d]) for r in rows] for d in range(3)])
    return out


def gpumd_inputs(run_in_path, xyz_path, mtp_path=None):
    """(model file, options for SUS2_SH, cell, pbc, types [N] int32, positions [3, N]):
    the species index of an atom is the position of its symbol on the potential line."""
    path, symbols, options = potential_line(run_in_path, mtp_path)
    xyz = read_xyz(xyz_path)
    index = {s: k for k, s in enumerate(symbols)}
    missing = sorted(set(xyz["symbols"]) - set(index))
    if missing:
        raise ValueError(f"element(s) {missing} of the model file are not on the potential line")
    types = np.array([index[s] for s in xyz["symbols"]], dtype=np.int32)
    return path, list(symbols) + list(options), xyz["cell"], xyz["pbc"], types, xyz["positions"]
