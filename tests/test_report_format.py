"""The report table is written in the reference's own text format
(write_qmc_report / write_qmc_report_header, src/qmc_io.f90:23-136,412-508,
733-803), so pyhande and HANDE users' scripts read it unchanged.  Checked
character for character against the header and first rows of the reference's
benchmark outputs (the values formatted are the ones those rows print)."""
import pytest

from hande_amd import report
from helpers import goldens

CASES = [k for k, e in goldens().items()
         if "report_text" in e and not any("RDM" in line for line in e["report_text"])]


def test_cases_cover_plain_restart_and_replica_free_tables():
    assert "He2-aug-cc-pVDZ_real_64_restart" in CASES and "hubbard_k_1d_n4" in CASES
    assert len(CASES) >= 10
    assert "He2-aug-cc-pVDZ_real_32_SS_replica" in CASES  # a two-space table


@pytest.mark.parametrize("name", CASES)
def test_report_lines_match_reference_text(name):
    e = goldens()[name]
    text = e["report_text"]
    nspaces = 2 if "Replica" in text[0] else 1  # a replica pair's table has a 'Replica i' line on top
    hdr = report.header(nspaces)
    assert hdr == text[:nspaces]
    # data rows only: bloom warnings (bloom_stats_warning) are comment lines without a leading number
    rows = [line for line in text[nspaces:] if line.lstrip(" #").split(" ")[0].isdigit()]
    assert len(rows) >= 3
    for i, (vals, line) in enumerate(zip(e["report"], rows)):
        assert report.row(vals, nspaces, comment=(i == 0)) == line


def test_replica_header_and_table():
    lines = report.header(2)
    assert lines[0].startswith(" #                  Replica 1") and "Replica 2" in lines[0]
    block = report.format_table([[0, 0.0, 0.0, 10.0, 10.0, 0.0, 0.0, 10.0, 10.0, 1, 0, 0.0, 0.0]], 2)
    assert block.startswith(report.PREAMBLE)
    assert block.rstrip("\n").splitlines()[-1].startswith(" #              0   0.0000000000E+00")
