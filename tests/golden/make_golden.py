"""Write the committed golden fixtures for the parity tests.

Run in the build container (needs the reference tree and the reference-source
dSFMT checker built by ``make -C oracle ref``):

    python tests/golden/make_golden.py

Outputs (small JSON data files, no reference source):
  * murmur_vectors.json -- MurmurHash2 values written by the reference's own C
    source (lib/external/MurmurHash2.c), the hash that assigns determinants to
    MPI ranks.
  * dsfmt_vectors.json -- close-open dSFMT-19937 stream values written by the
    reference's own C source (lib/dSFMT-src-2.2.3/dSFMT.c) with the 50000-double
    refill of lib/dSFMT_F03_interface/dSFMT_interface.F90:387-409.
  * qmc_goldens.json -- for each reference regression test on the hot path:
    the input deck (test input data), the QMC report table, and the setup
    tables the reference printed (basis functions, momentum symmetry table,
    reference determinant).  Row selection and tolerance follow
    tools/pyhande/pyhande/testcode.py:57-61 and test_suite/userconfig:14-15.
"""
import gzip
import hashlib
import json
import os
import re
import struct
import subprocess
import sys

REF = os.environ.get("HANDE_REFERENCE", "/root/reference")
HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
DSFMT_REF = os.path.join(ROOT, "oracle", "_ref", "dsfmt_ref")
MURMUR_REF = os.path.join(ROOT, "oracle", "_ref", "murmur_ref")

QMC_TESTS = [
    # (name, directory under test_suite, input file)
    ("hubbard_k_1d_n4", "fciqmc/np1/hubbard_k_1d_n4", "hubbard.in"),
    ("hubbard_k_nontilted_2d_fciqmc", "fciqmc/np1/hubbard_k_nontilted_2d_fciqmc", "hubbard.in"),
    ("hubbard_k_1d_renorm", "fciqmc/np1/hubbard_k_1d_no_renorm", "renorm.in"),
    ("hubbard_k_1d_no_renorm", "fciqmc/np1/hubbard_k_1d_no_renorm", "no_renorm.in"),
    ("hubbard_k_1d_ifciqmc", "ifciqmc/np1/hubbard_k_1d_ifciqmc", "hubbard.in"),
    ("hubbard_k_nontilted_2d_ifciqmc", "ifciqmc/np1/hubbard_k_nontilted_2d_ifciqmc", "hubbard.in"),
    ("hubbard_k_1d_ifciqmc_cas", "ifciqmc/np1/hubbard_k_1d_ifciqmc_cas", "hubbard.in"),
    ("hubbard_k_1d_ifciqmc_real_32", "ifciqmc_real_32/np1/hubbard_k_1d_ifciqmc_real_32", "hubbard.in"),
    ("H2O-RHF-cc-pVTZ", "fciqmc/np1/H2O-RHF-cc-pVTZ", "h2o.in"),
    ("Ne-RHF-aug-cc-pVDZ", "ifciqmc/np1/Ne-RHF-aug-cc-pVDZ", "hande_dump.in"),
    ("H2-RHF-cc-pVTZ-Lz", "fciqmc/np1/H2-RHF-cc-pVTZ-Lz", "Lz.in"),
    ("H2-sto-3g-replica", "fciqmc/np1/H2-sto-3g-replica", "h2.in"),
    # two MPI ranks (real_64 populations, initiator): replayed on emulated ranks
    ("hubbard_k_14s_real_64_np2", "ifciqmc_real_64/np2/hubbard_k_14s_14e_U2_real_64", "hubbard.in"),
    # a run continued from the reference's restart file (HDF5, integer populations
    # read into real_64 amplitudes)
    ("He2-aug-cc-pVDZ_real_64_restart", "fciqmc_real_64/np1/He2-aug-cc-pVDZ_real_64_restart", "restart.in"),
    # restart + initiator + vary_shift from the start + on-the-fly blocking with
    # automatic shift damping
    ("He2-aug-cc-pVDZ_real_64_auto_shift_damping_restart",
     "ifciqmc_real_64/np1/He2-aug-cc-pVDZ_real_64_auto_shift_damping_restart", "read.in"),
    # semi-stochastic projection (space "high", core chosen at iteration 1000):
    # combined annihilation, then separate annihilation with a replica pair
    ("He2-aug-cc-pVDZ_real_32_SS", "fciqmc_real_32/np1/He2-aug-cc-pVDZ_real_32_SS", "he2.in"),
    ("He2-aug-cc-pVDZ_real_32_SS_replica", "fciqmc_real_32/np1/He2-aug-cc-pVDZ_real_32_SS_replica", "he2.in"),
]

# long tables are stored truncated (the tests compare a prefix)
MAX_ROWS = {"H2-sto-3g-replica": 2001}

FCI_TESTS = [
    ("fci_hubbard_k_1d", "fci/hubbard_k_1d", "hubbard.in"),
    ("fci_hubbard_k_nontilted_2d", "fci/hubbard_k_nontilted_2d_fci", "hubbard.in"),
    ("fci_hubbard_k_tilted_2d", "fci/hubbard_k_tilted_2d_fci", "hubbard.in"),
    ("fci_H2O-RHF-sto-2g", "fci/H2O-RHF-sto-2g", "h2o_fci.in"),
]

NUM = r"[-+]?\d+\.?\d*(?:[EeDd][-+]?\d+)?"


def dsfmt_vectors():
    out = {"source": "lib/dSFMT-src-2.2.3/dSFMT.c via oracle/ref_dsfmt_driver.c",
           "refill": 50000, "streams": []}
    for seed in (7, 8, 9, 5234, -538938605, 1233392714):
        n = 120000
        txt = subprocess.run([DSFMT_REF, str(seed), str(n)], check=True,
                             capture_output=True, text=True).stdout.split()
        assert len(txt) == n
        out["streams"].append({
            "seed": seed,
            "count": n,
            "first": txt[:64],
            # values straddling the first and second 50000-element refills
            "around_refill": {str(i): txt[i] for i in
                              list(range(49995, 50005)) + list(range(99995, 100005))},
            "last": txt[-8:],
            "sha256": hashlib.sha256("\n".join(txt).encode()).hexdigest(),
        })
    return out


def murmur_vectors():
    """MurmurHash2 of byte strings written by the reference's own C source
    (lib/external/MurmurHash2.c): every length 0..24 and determinant-like
    4-, 8- and 16-byte strings, seeds 7 (the rank hash) and others."""
    import random
    rnd = random.Random(2024)
    cases = []
    for seed in (7, 0, 1, 0xdeadbeef):
        for n in range(25):
            cases.append((seed, bytes(rnd.getrandbits(8) for _ in range(n))))
    for _ in range(200):
        n = rnd.choice([4, 8, 16])
        cases.append((7, bytes(rnd.getrandbits(8) for _ in range(n))))
    inp = "".join(f"{sd} {b.hex()}\n" if b else f"{sd}\n" for sd, b in cases)
    txt = subprocess.run([MURMUR_REF], input=inp, check=True, capture_output=True, text=True).stdout.split()
    assert len(txt) == len(cases)
    return {"source": "lib/external/MurmurHash2.c via oracle/ref_murmur_driver.c",
            "cases": [{"seed": sd, "bytes": b.hex(), "hash": int(h)} for (sd, b), h in zip(cases, txt)]}


def parse_report_table(lines):
    """QMC report table: header line starts with '#     iterations'."""
    rows = []
    header = None
    for i, line in enumerate(lines):
        if re.match(r"^\s*#\s+iterations\s+Shift", line):
            # column names are separated by two or more spaces
            header = re.split(r"\s{2,}", line.strip().lstrip("#").strip())
            start = i + 1
            break
    if header is None:
        return None
    # density-matrix columns (replica runs with density_matrices) are not
    # produced by the walker-propagation path: drop them
    drop = {k for k, name in enumerate(header) if name.startswith("RDM")}
    for line in lines[start:]:
        s = line.strip()
        if not s:
            break
        if s.startswith("#"):
            s = s.lstrip("#").strip()
            # the iteration-0 row is printed as a comment row
        parts = s.split()
        if not re.fullmatch(r"\d+", parts[0]):
            continue
        rows.append([float(x) for k, x in enumerate(parts) if k not in drop])
    return rows


def parse_shift_damping_changes(lines):
    """(iterations of the report row above, new damping) for every
    '# Auto-Shift-Damping: Shift damping changed to' line of the table."""
    out, last = [], None
    for line in lines:
        m = re.match(r"^\s*#?\s*(\d+)\s+" + NUM, line)
        if m:
            last = int(m.group(1))
        m = re.search(r"Shift damping changed to\s+(" + NUM + ")", line)
        if m and last is not None:
            out.append([last, float(m.group(1))])
    return out


def parse_basis(lines):
    basis = []
    for i, line in enumerate(lines):
        if line.strip().startswith("index") and ("<i|f|i>" in line or "sym" in line):
            for l2 in lines[i + 1:]:
                if not l2.strip():
                    if basis:
                        return basis
                    continue
                m = re.match(r"^\s*(\d+)\s+(.*)$", l2)
                if not m:
                    return basis
                basis.append(l2.rstrip())
            return basis
    return basis


def parse_symmetry(lines):
    out = {}
    for i, line in enumerate(lines):
        if "The table below gives the label and inverse of each wavevector" in line:
            inv = []
            for l2 in lines[i + 3:]:
                if not l2.strip():
                    break
                inv.append(int(l2.split()[-1]))
            out["inv_sym"] = inv
        if "The matrix below gives the result of k_i+k_j" in line:
            tab = []
            for l2 in lines[i + 2:]:
                if not l2.strip():
                    break
                tab.append([int(x) for x in l2.split()])
            out["sym_table_printed"] = tab  # row i, column j holds sym_table(j,i)
    return out


def parse_json_block(text):
    """Pick the scalar keys the parity tests need out of the reference's JSON echo.

    The echo is not strict JSON (Fortran formatting), so keys are read by pattern."""
    m = re.search(r"-- Start JSON block --\n(.*?)\n -- End JSON block --", text, re.S)
    if not m:
        return None
    body = m.group(1)
    out = {}
    for key in ("nbasis", "nel", "nalpha", "nbeta", "nvirt_alpha", "nvirt_beta", "symmetry",
                "rng_seed", "tau", "shift_damping", "initial_shift", "target_particles",
                "initiator_pop", "spawn_cutoff", "pattempt_single", "pattempt_double",
                "pattempt_parallel", "H00", "F0", "det_symmetry", "U", "t", "ncycles", "nreport",
                "walker_length", "spawned_walker_length", "D0_population"):
        mm = re.search(r'"%s":\s*(%s)' % (re.escape(key), NUM), body)
        if mm:
            out[key] = float(mm.group(1).replace("D", "E"))
    for key in ("real_amplitudes", "real_amplitude_force_32", "initiator_approx", "replica_tricks",
                "quadrature_initiator"):
        mm = re.search(r'"%s":\s*(true|false)' % key, body)
        if mm:
            out[key] = mm.group(1) == "true"
    mm = re.search(r'"excit_gen":\s*"(\w+)"', body)
    if mm:
        out["excit_gen"] = mm.group(1)
    mm = re.search(r'"det":\s*\[([^\]]*)\]', body)
    if mm:
        out["det"] = [int(x) for x in mm.group(1).split(",")]
    return out


def main():
    if not os.path.exists(DSFMT_REF):
        sys.exit("build oracle/_ref/dsfmt_ref first (make -C oracle ref)")
    with open(os.path.join(HERE, "dsfmt_vectors.json"), "w") as f:
        json.dump(dsfmt_vectors(), f, indent=1)
    with open(os.path.join(HERE, "murmur_vectors.json"), "w") as f:
        json.dump(murmur_vectors(), f, indent=1)

    goldens = {}
    for name, d, inp in QMC_TESTS + FCI_TESTS:
        path = os.path.join(REF, "test_suite", d)
        bench = [x for x in os.listdir(path) if x.startswith("benchmark.out") and x.endswith("inp=" + inp)]
        assert len(bench) == 1, (d, bench)
        text = open(os.path.join(path, bench[0])).read()
        lines = text.splitlines()
        entry = {
            "test": f"test_suite/{d}/{bench[0]}",
            "input": open(os.path.join(path, inp)).read(),
            "basis": parse_basis(lines),
        }
        entry.update(parse_symmetry(lines))
        m = re.search(r"Number of MPI processes running on:\s*(\d+)", text)
        entry["nprocs"] = int(m.group(1)) if m else 1
        js = parse_json_block(text)
        if js is not None:
            entry["json"] = js
        table = parse_report_table(lines)
        if table is not None:
            entry["report"] = table[:MAX_ROWS.get(name, len(table))]
            # the header and first rows verbatim, for the text-format check
            hdr = next(i for i, line in enumerate(lines) if re.match(r"^\s*#\s+iterations\s+Shift", line))
            entry["report_text"] = lines[hdr - (1 if "Replica" in lines[hdr - 1] else 0):hdr + 6]
            changes = parse_shift_damping_changes(lines)
            if changes:
                entry["shift_damping_changes"] = changes
        if name.startswith("fci_"):
            eig = []
            for i, line in enumerate(lines):
                if re.match(r"^\s*State\s+Energy", line):
                    for l2 in lines[i + 1:]:
                        if not l2.strip():
                            break
                        eig.append(float(l2.split()[1]))
                    break
            entry["eigenvalues"] = eig
        # data files the deck reads (integral dumps, restart files), stored gzipped
        entry["data_files"] = []
        for x in sorted(os.listdir(path)):
            is_restart = (re.fullmatch(r"HANDE\.RS\.\d+\.p\d+\.H5", x) is not None
                          and re.search(r"restart\s*=\s*\{[^}]*read", entry["input"]) is not None)
            if is_restart or (x in ("FCIDUMP", "INTDUMP") and not name.startswith("fci_")
                              and not name.endswith("-Lz")):
                # the integral dump (or restart file) the deck reads: test input data, stored gzipped
                dst = os.path.join(HERE, f"{name}.{x}.gz")
                with open(os.path.join(path, x), "rb") as src, open(dst, "wb") as raw, \
                        gzip.GzipFile(filename="", mode="wb", compresslevel=9, fileobj=raw, mtime=0) as out:
                    out.write(src.read())
                entry["data_files"].append(os.path.basename(dst))
        goldens[name] = entry
    with open(os.path.join(HERE, "qmc_goldens.json"), "w") as f:
        json.dump(goldens, f, indent=1)
    print("wrote", len(goldens), "golden entries")


if __name__ == "__main__":
    main()
