#!/usr/bin/env python3
"""Writes the golden fixtures of tests/golden (run once, here, where oracle/_ref exists).

For each parameter set, from the secret key shipped with the reference
(kem/test_vectors/Cipher_Sk/sk_<p>.bin, kept here as sk_<p>.bin):
  L, g     support and Goppa polynomial from the reference CPU decoder's own support_gen /
           key parsing (oracle/_ref/libref_<p>.so, built from the reference's C sources);
  ct, e    5 ciphertexts made by this library's seeded generator and their error vectors;
  stream   the errorstream0_<p>.bin the reference driver (oracle/_ref/bin/decrypt_<p>)
           writes for those ciphertexts.
Plus golden_c1_348864.npz: the CPU-reference configuration (mceliece348864, one seeded
key, 5 ciphertexts of weight 64) with the reference driver's errorstream.
"""
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

import cm_decrypt__amd as cm  # noqa: E402
import _oracle as orc  # noqa: E402


def main():
    for pid in cm.PARAM_IDS:
        sk = orc.shipped_sk(pid)
        ref = orc.Ref(pid)
        ref.set_key(sk)
        g, L, _ = ref.key()
        ct, e = cm.gen_encrypt(pid, sk, 4242, 5)
        with tempfile.TemporaryDirectory() as d:
            rc, _, stream = orc.run_ref_driver(pid, sk, ct, d)
        assert rc == 0 and stream is not None
        np.savez_compressed(os.path.join(HERE, f"golden_{pid}.npz"), L=L, g=g, ct=ct, e=e,
                            stream=np.frombuffer(stream, np.uint8))
        print(pid, "ok", len(stream), "errorstream bytes")
    pid = 348864
    sk = cm.gen_keypair(pid, 1)
    ct, e = cm.gen_encrypt(pid, sk, 1, 5)
    with tempfile.TemporaryDirectory() as d:
        rc, _, stream = orc.run_ref_driver(pid, sk, ct, d)
    assert rc == 0
    np.savez_compressed(os.path.join(HERE, "golden_c1_348864.npz"), sk=sk, ct=ct, e=e,
                        stream=np.frombuffer(stream, np.uint8))
    print("c1 ok")


if __name__ == "__main__":
    main()
