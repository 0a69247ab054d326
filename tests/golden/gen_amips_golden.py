"""Writes tests/golden/amips_golden.npz: inputs and the outputs of the
reference's own comformalAMIPSEnergy_new / Jacobian_new / Hessian_new
(src/tetwild/LocalOperations.cpp:28-291), compiled from the reference sources
by oracle/ref_build.sh.

Run in the container that holds /root/reference:
    python tests/golden/gen_amips_golden.py
The fixture holds data only (inputs, outputs, a kind code per row); the
reference source never leaves this container.

Kinds: 0 regular tets (random similarity transforms), 1 jittered Kuhn-lattice
tets (jitter N(0, 0.15h)), 2 aspect ratio 10^U[0,4] pancakes/needles,
3 inverted tets (two vertices swapped), 4 exactly degenerate tets (det = 0),
5 Gaussian random tets.
"""
import json
import os
import subprocess
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from tetwild_amd import workloads as W  # noqa: E402


def gather(tets, x, y, z):
    return np.stack([np.stack([x[tets[:, j]], y[tets[:, j]], z[tets[:, j]]], 1)
                     for j in range(4)], 1).reshape(-1, 12)


def main():
    if O.ref() is None:
        subprocess.run([os.path.join(ROOT, "oracle", "ref_build.sh")], check=True)
    st = W.Stream(2024, 1)
    rows, kinds = [], []

    # 0: regular tets under random similarity transforms
    R = W._rotations(st, 64)
    sc = 10.0 ** st.uniform(64, -3, 1)
    tr = st.uniform(64 * 3, -5, 5).reshape(64, 3)
    P = np.einsum("nij,vj->nvi", R, W.REGULAR_TET) * sc[:, None, None] + tr[:, None, :]
    rows.append(P.reshape(-1, 12)); kinds += [0] * 64

    # 1: jittered Kuhn lattice
    x, y, z, t, _ = W.kuhn_lattice(6, seed=7)
    sel = st.integers(600, 0, len(t) - 1)
    rows.append(gather(t[sel], x, y, z)); kinds += [1] * 600

    # 2: aspect ratio up to 1e4
    x, y, z, t = W.aspect_tets(600, seed=9)
    rows.append(gather(t, x, y, z)); kinds += [2] * 600

    # 3: inverted (swap vertices 2 and 3 of lattice tets)
    x, y, z, t, _ = W.kuhn_lattice(5, seed=11)
    sel = st.integers(200, 0, len(t) - 1)
    ti = t[sel][:, [0, 1, 3, 2]]
    rows.append(gather(ti, x, y, z)); kinds += [3] * 200

    # 4: exactly degenerate: small-integer coordinates, v3 = v0 + (v1-v0) + (v2-v0)
    a = st.integers(50 * 9, -20, 20).reshape(50, 3, 3).astype(np.float64)
    v3 = a[:, 1] + a[:, 2] - a[:, 0]
    D = np.concatenate([a, v3[:, None, :]], 1).reshape(-1, 12)
    rows.append(D); kinds += [4] * 50

    # 5: Gaussian tets
    rows.append(st.normal(400 * 12).reshape(400, 12)); kinds += [5] * 400

    T = np.ascontiguousarray(np.concatenate(rows))
    kinds = np.array(kinds, dtype=np.int8)
    E, G, H = O.ref_ejh_T(T)

    # timing: compiled reference vs oracle restatement, one core
    x, y, z, t, _ = W.kuhn_lattice(40, seed=1)
    TT = np.ascontiguousarray(gather(t, x, y, z))
    n = TT.shape[0]
    E2 = np.empty(n); G2 = np.empty((n, 3)); H2 = np.empty((n, 9))
    t0 = time.perf_counter()
    O.ref().ref_amips_batch_ejh(O._p(TT), n, O._p(E2), O._p(G2), O._p(H2))
    t_ref = time.perf_counter() - t0
    Es = np.empty(n); Gs = np.empty((n, 3)); Hs = np.empty((n, 9))
    t0 = time.perf_counter()
    O.lib().orc_amips_batch(O._p(np.ascontiguousarray(t)), n, O._p(x), O._p(y), O._p(z), 2,
                            O._p(Es), O._p(Gs), O._p(Hs))
    t_orc = time.perf_counter() - t0

    np.savez_compressed(os.path.join(HERE, "amips_golden.npz"), T=T, kind=kinds, E=E, G=G, H=H)
    meta = {
        "source": "src/tetwild/LocalOperations.cpp:28-291 compiled by oracle/ref_build.sh "
                  "(g++ -O3 -ffp-contract=off)",
        "rows": int(T.shape[0]),
        "timing_tets": int(n),
        "ref_ejh_tets_per_s_1core": n / t_ref,
        "oracle_ejh_tets_per_s_1core": n / t_orc,
        "oracle_over_ref_speed": t_ref / t_orc,
    }
    with open(os.path.join(HERE, "amips_golden_meta.json"), "w") as f:
        json.dump(meta, f, indent=1)
    print(json.dumps(meta, indent=1))


if __name__ == "__main__":
    main()
