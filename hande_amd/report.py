"""The QMC report table as HANDE prints it, so its own analysis tools read it.

HANDE writes one line per report (write_qmc_report, src/qmc_io.f90:412-508)
under a column header (write_qmc_report_header, src/qmc_io.f90:23-136), with
the field formats of write_qmc_var (src/qmc_io.f90:733-803): integers as
``i14`` (64-bit ones ``i20``), reals as ``es17.10`` followed by four spaces,
R_spawn and the time per cycle as ``f8.4``; the first (iteration-0) line is a
comment line.  Column titles follow write_column_title (src/qmc_io.f90:354-410).
pyhande's extractor (tools/pyhande/pyhande/extract.py) finds the table by its
header line and reads the rows as whitespace-separated numbers.
"""
from __future__ import annotations

import numpy as np

PREAMBLE = (
    " Information printed out every QMC report loop:\n\n"
    " Shift: the energy offset calculated at the end of the report loop.\n"
    " H_0j: <D_0|H|D_j>, Hamiltonian matrix element.\n"
    " N_j: population of Hamiltonian particles on determinant D_j.\n"
    " # H psips: current total population of Hamiltonian particles.\n"
    " # states: number of many-particle states occupied.\n"
    " spawn_events: number of successful spawning events across all processors.\n"
    " R_spawn: average rate of spawning across all processors.\n"
    " time: average time per Monte Carlo cycle.\n\n"
    " Note that all particle populations are averaged over the report loop.\n"
)


def _title(key, kind="real", justify=0):
    """write_column_title: a 20-character key cut to the field's width."""
    width = {"int": 14, "int64": 20, "low": 8}.get(kind, 20)
    k = key.ljust(20)[:width]
    if justify == 1:  # right justify
        return k.rstrip().rjust(width) + "  "
    if justify == 2:  # as given
        return k + "  "
    return " " + k[:width - 1].strip().ljust(width - 1) + "  "


def _int(v):
    return f"{int(round(v)):14d}  "


def _real(v):
    s = f"{float(v):17.10E}"
    return s + "     "


def _low(v):
    return f"{float(v):8.4f}  "


def header(nspaces=1):
    """The table's header line(s) (a 'Replica i' line above when nspaces > 1)."""
    lines = []
    if nspaces > 1:
        line = " # " + _title("", "int")
        for i in range(nspaces):
            line += _title(f"Replica {i + 1}") + _title("") * 3
        lines.append(line)
    line = " # " + _title("iterations", "int", justify=1)
    for _ in range(nspaces):
        line += _title("Shift") + _title("\\sum H_0j N_j") + _title("N_0") + _title("# H psips")
    line += _title("# states", "int", 1) + _title("# spawn_events", "int", 1)
    line += _title("R_spawn", "low") + _title("  time", "low", 2)
    lines.append(line)
    return lines


def row(values, nspaces=1, comment=False):
    """One report line from a row of ``api.rows_to_table`` (iterations, then
    shift, sum H_0j N_j, N_0, # H psips per space, then # states,
    # spawn_events, R_spawn) and, if present, the time per cycle."""
    v = list(values)
    line = (" # " if comment else "   ") + _int(v[0])
    k = 1
    for _ in range(nspaces):
        line += "".join(_real(x) for x in v[k:k + 4])
        k += 4
    line += _int(v[k]) + _int(v[k + 1]) + _low(v[k + 2])
    line += _low(v[k + 3] if len(v) > k + 3 else 0.0)
    return line


def format_table(table, nspaces=1, preamble=True):
    """The whole report block: preamble, header, then one line per row (the
    first as a comment line, as initial_qmc_status prints it)."""
    out = [PREAMBLE] if preamble else []
    out += header(nspaces)
    for i, r in enumerate(np.asarray(table)):
        out.append(row(r, nspaces, comment=(i == 0)))
    return "\n".join(out) + "\n"
