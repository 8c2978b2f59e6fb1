"""Write tests/golden/ref_small.npz from the compiled reference (oracle/_ref).

The fixture holds inputs and the reference's outputs only (no source):
a small integer mixture, the reference's HNSW-seeded KNN lists, its refined
lists after refineKNNG run with one thread (deterministic), its reverse CSR,
phase-A proxies and searchRknn results for a few queries, and L2Sqr values
on float pairs.  Run from the repository root:

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import hrnn_amd  # noqa: E402
import oracle  # noqa: E402


def main():
    n, nq, d, K, iters, S = 1500, 24, 64, 24, 2, 8
    m, k, Kp, ef = 6, 5, 24, 32
    x = hrnn_amd.synth_mixture(41, n, d, 20, 20.0, 0.0, 255.0, True, 0.0, 1041)
    q = hrnn_amd.synth_mixture(42, nq, d, 20, 20.0, 0.0, 255.0, True, 0.0, 1041)
    r = oracle.RefIndex(n, d, M=8, efc=32, K=K)
    r.add_points(x, nthreads=1)
    sid, sd, sc = r.get_knng()
    r.refine(iters, S, nthreads=1)
    rid, rd, rc = r.get_knng()
    r.build_rknng()
    off, ent, rb, wide = r.get_csr()
    assert wide == 0
    prox = r.proxies(q, m, ef)
    roff, rlab = r.search(q, m, 1.0, k, Kp, ef)
    rng = np.random.default_rng(43)
    a = rng.standard_normal((64, 128)).astype(np.float32)
    b = rng.standard_normal((64, 128)).astype(np.float32)
    r2 = oracle.RefIndex(4, 128, K=0)
    l2 = np.array([r2.l2(a[i], b[i]) for i in range(64)], np.float32)
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ref_small.npz")
    np.savez_compressed(out, x=x, q=q, K=K, iters=iters, sample=S, seed_ids=sid, seed_d=sd,
                        seed_cnt=sc, ref_ids=rid, ref_d=rd, ref_cnt=rc, csr_off=off, csr_ent=ent,
                        proxies=prox, k=k, Kp=Kp, res_off=roff, res_lab=rlab, l2_a=a, l2_b=b,
                        l2_ref=l2)
    print("wrote", out, os.path.getsize(out), "bytes")


if __name__ == "__main__":
    main()
