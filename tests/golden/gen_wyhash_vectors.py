"""Write tests/golden/wyhash_vectors.json from the reference's own wyhash.h.

Runs in a container that holds the reference tree: `make -C oracle ref` first
builds oracle/_ref/libwyref.so from the reference's include/wyhash/wyhash.h
(with Cuttlefish's salt, include/Minimizer_Utility.hpp:26).  The vectors pin
the oracle's restatement of wyhash (every length branch of wyhash() and the
8-byte l-mer hash Minimizer_Utility::hash uses) and, through the oracle, the
HIP l-mer hash.
"""
import ctypes
import json
import os
import random

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))


def main():
    lib = ctypes.CDLL(os.path.join(REPO, "oracle", "_ref", "libwyref.so"))
    lib.ref_wyhash.restype = ctypes.c_uint64
    lib.ref_wyhash.argtypes = [ctypes.c_char_p, ctypes.c_size_t, ctypes.c_uint64]
    lib.ref_lmer_hash.restype = ctypes.c_uint64
    lib.ref_lmer_hash.argtypes = [ctypes.c_uint64, ctypes.c_uint64]

    rng = random.Random(20261203)
    lmers = [0, 1, 2, 3, (1 << 40) - 1, (1 << 64) - 1, 0x5555555555, 0xAAAAAAAAAA]
    for l in (5, 15, 20, 27, 32):
        for _ in range(40):
            lmers.append(rng.getrandbits(2 * l))
    lmer_vecs = [[str(x), str(lib.ref_lmer_hash(x, 0))] for x in lmers]

    byte_vecs = []
    for n in list(range(0, 70)) + [95, 96, 97, 128, 200]:
        for seed in (0, 0xAAAAAAAA55555555):
            data = bytes(rng.getrandbits(8) for _ in range(n))
            byte_vecs.append([data.hex(), str(seed), str(lib.ref_wyhash(data, n, seed))])

    out = {
        "source": "reference include/wyhash/wyhash.h (v4.2) with Cuttlefish salt",
        "lmer_hash_seed0": lmer_vecs,
        "wyhash_bytes": byte_vecs,
    }
    with open(os.path.join(HERE, "wyhash_vectors.json"), "w") as f:
        json.dump(out, f, indent=0)
    print("wrote", len(lmer_vecs), "l-mer vectors and", len(byte_vecs), "byte vectors")


if __name__ == "__main__":
    main()
