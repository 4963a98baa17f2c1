"""Seeded synthetic molecular Hamiltonians in FCIDUMP format.

The benchmark systems are molecular-like Hamiltonians built from random
integrals (no chemistry package is available): NORB spatial orbitals in C1
symmetry with
  * h_pq symmetric, diagonal eps_p ascending on [-2, 3] Ha, off-diagonal
    N(0, 0.05) damped by exp(-|p-q|/4);
  * (pq|rs) = sum_P L^P_pq L^P_rs over NCHOL vectors with L^P_pq ~ N(0, 0.1)
    damped by exp(-|p-q|/4) (positive semi-definite, 8-fold symmetric);
  * the orbital energies eps_p written as "eps p 0 0 0" lines, which the
    integral reader uses to order the orbitals (src/read_in.F90:1175-1178);
  * every value written with 17 significant digits so the file round-trips.
"""
import numpy as np


def synthetic_integrals(norb, seed, nchol=80, chol_scale=0.1):
    """(eps, h, eri) of the recipe above; chol_scale is the standard deviation of
    the Cholesky vectors (0.1 for the benchmark systems; smaller values give
    weakly correlated test systems whose ground state is close to the
    reference determinant)."""
    rng = np.random.default_rng(seed)
    eps = np.linspace(-2.0, 3.0, norb)
    idx = np.arange(norb)
    damp = np.exp(-np.abs(idx[:, None] - idx[None, :]) / 4.0)
    h = rng.normal(0.0, 0.05, (norb, norb)) * damp
    h = np.triu(h, 1)
    h = h + h.T + np.diag(eps)
    L = rng.normal(0.0, chol_scale, (nchol, norb, norb)) * damp[None]
    L = 0.5 * (L + L.transpose(0, 2, 1))
    eri = np.einsum("Ppq,Prs->pqrs", L, L, optimize=True)
    return eps, h, eri


def write_fcidump(path, norb, nelec, seed, nchol=80, ecore=0.0, chol_scale=0.1):
    eps, h, eri = synthetic_integrals(norb, seed, nchol, chol_scale)
    lines = [f" &FCI NORB={norb},NELEC={nelec},MS2=0,",
             "  ORBSYM=" + ",".join(["1"] * norb) + ",",
             "  ISYM=1,", " &END"]
    fmt = "{:.16e} {:d} {:d} {:d} {:d}"
    for p in range(norb):
        for q in range(p + 1):
            pq = p * (p + 1) // 2 + q
            for r in range(norb):
                for s in range(r + 1):
                    rs = r * (r + 1) // 2 + s
                    if rs > pq:
                        continue
                    lines.append(fmt.format(eri[p, q, r, s], p + 1, q + 1, r + 1, s + 1))
    for p in range(norb):
        for q in range(p + 1):
            lines.append(fmt.format(h[p, q], p + 1, q + 1, 0, 0))
    for p in range(norb):
        lines.append(fmt.format(eps[p], p + 1, 0, 0, 0))
    lines.append(fmt.format(ecore, 0, 0, 0, 0))
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
    return path


# Benchmark configurations: (norb, nelec, seed)
MOLECULAR_CONFIGS = {
    "mol26": (26, 10, 26010),
    "mol40": (40, 14, 40014),
    "mol60": (60, 20, 60020),
}
