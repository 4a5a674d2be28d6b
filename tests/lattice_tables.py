"""The oracle's lattice tables (oracle/oracle_lattice.h) read back through a
tiny C probe, for tests that restate per-node sums in numpy."""
import os
import subprocess
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _tables(q):
    """Tables of oracle/oracle_lattice.h read back through a tiny C probe."""
    src = r"""
#include <stdio.h>
#include "oracle_lattice.h"
int main(int argc, char **argv) {
    int q = atoi(argv[1]);
    const oracle_lattice *L = oracle_lattice_for(q);
    for (int s = 0; s < q; ++s)
        printf("%d %d %d %d %.17g %.17g\n", L->ex[s], L->ey[s], L->ez[s], L->bb[s],
               L->w_is_float ? (double)L->wf[s] : L->wd[s], L->wd[s]);
    printf("UNK");
    for (int i = 0; i < L->n_unk; ++i) printf(" %d %d %d %d %d %d", L->in_unk[i], L->in_par[i],
        L->out_unk[i], L->out_par[i], L->in_list[i], L->out_list[i]);
    printf("\nZERO");
    for (int i = 0; i < L->n_zero; ++i) printf(" %d", L->zero_list[i]);
    printf("\n");
    return 0;
}
"""
    return _run_c_probe(src, ["-I", os.path.join(ROOT, "oracle")], [str(q)], "#include <stdlib.h>\n")


def _run_c_probe(src, flags, args, prelude="", compiler="gcc"):
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "probe.c" if compiler == "gcc" else "probe.cpp")
        with open(c, "w") as fh:
            fh.write(prelude + src)
        exe = os.path.join(d, "probe")
        subprocess.run([compiler, "-O1", *flags, c, "-o", exe], check=True)
        return subprocess.run([exe, *args], check=True, capture_output=True, text=True).stdout


def _parse_tables(text):
    lines = text.strip().splitlines()
    rows = [ln.split() for ln in lines if not ln.startswith(("UNK", "ZERO"))]
    e = np.array([[int(r[0]), int(r[1]), int(r[2])] for r in rows])
    bb = np.array([int(r[3]) for r in rows])
    w = np.array([float(r[4]) for r in rows])
    w_exact = np.array([float(r[5]) for r in rows])
    unk = np.array([int(v) for v in lines[-2].split()[1:]]).reshape(-1, 6)
    zero = [int(v) for v in lines[-1].split()[1:]]
    return e, bb, w, w_exact, unk, zero


def tables(q):
    """e (Q,3), bb, reference weights, exact weights of the D3Qq lattice."""
    e, bb, w, w_exact, _, _ = _parse_tables(_tables(q))
    return e, bb, w, w_exact
