"""Host-side file formats of the KMC path: POSCAR input and `*.log` reports.

`read_poscar` accepts the VASP-style layout read by PyCD/io.py:12-82: an
optional element-symbol line 6, lattice rows on lines 3-5 (the scale line 2
is not applied, as in the reference), counts, a `Direct`/`Cartesian` tag and
one coordinate row per site.

`generate_report` writes `<name>.log` with the same text layout as
PyCD/io.py:163-199.  The layout is a file contract: the reference's run
driver parses alpha from characters [7:16] of line 4 of
`precomputed_array.log` (PyCD/material_run.py:75-80).
"""
from datetime import datetime

import numpy as np


def read_poscar(path):
    with open(path, "r") as fh:
        lines = fh.read().splitlines()
    first_tokens = lines[5].split()
    has_symbols = bool(first_tokens) and first_tokens[0].isalpha()
    lattice = np.array([[float(v) for v in lines[i].split()[:3]]
                        for i in range(2, 5)], dtype=np.float64)
    if has_symbols:
        element_types = lines[5].split()
        count_line = 6
    else:
        element_types = lines[0].split()
        count_line = 5
    counts = np.array([int(v) for v in lines[count_line].split()], dtype=int)
    coordinate_type = lines[count_line + 1].split()[0]
    total = int(counts.sum())
    coords = np.zeros((total, 3))
    first = count_line + 2
    for k in range(total):
        coords[k] = [float(v) for v in lines[first + k].split()[:3]]
    width = len(lines[first].split()[0])
    file_format = {18: "VASP", 11: "VESTA"}.get(width, "unknown")
    return {
        "lattice_matrix": lattice,
        "element_types": element_types,
        "num_elements": counts,
        "total_elements": total,
        "coordinate_type": coordinate_type,
        "coordinates": coords,
        "file_format": file_format,
    }


def generate_report(start_time, dst_path, file_name, print_time_elapsed,
                    prefix=None):
    elapsed = datetime.now() - start_time
    text = prefix if prefix else ""
    if print_time_elapsed:
        secs = elapsed.seconds
        days = f"{elapsed.days:2d} days, " if elapsed.days else ""
        text += (f"Time elapsed: {days}{(secs // 3600) % 24:2d} hours"
                 f", {(secs // 60) % 60:2d} minutes, {secs % 60:2d} seconds")
    with open(dst_path / (file_name + ".log"), "w") as fh:
        fh.write(text)
