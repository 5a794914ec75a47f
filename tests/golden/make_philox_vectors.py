"""Write tests/golden/philox2x32_10.json: outputs of the reference's own
random123() (include/bohrium/jitk/kernel_dependencies/random123_openmp.h +
thirdparty/Random123-1.09 philox.h), compiled from the reference tree by
`make -C oracle ref` into oracle/_ref/libr123_ref.so.

Run in the development container (where /root/reference exists):
    make -C oracle ref && python tests/golden/make_philox_vectors.py
"""
import ctypes
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "..", "..", "oracle", "_ref", "libr123_ref.so")


def main():
    lib = ctypes.CDLL(LIB)
    lib.ref_random123.restype = ctypes.c_uint64
    lib.ref_random123.argtypes = [ctypes.c_uint64] * 3
    cases = []

    def add(name, start, key, indices):
        out = [int(lib.ref_random123(start, key, int(i))) for i in indices]
        cases.append({"name": name, "start": int(start), "key": int(key),
                      "index": [int(i) for i in indices], "out": [str(o) for o in out]})

    add("key42_start0_first64", 0, 42, range(64))
    add("key42_start0_sparse", 0, 42, [10 ** k for k in range(8)] + [10 ** 7 - 1])
    add("key42_start1e7_first64", 10 ** 7, 42, range(64))
    add("key42_start1e7_tail", 10 ** 7, 42, [10 ** 7 - 3, 10 ** 7 - 2, 10 ** 7 - 1])
    add("counter_crosses_2pow32", 2 ** 32 - 4, 7, range(8))
    add("counter_high_word", 2 ** 40 + 12345, 99, range(8))
    add("key_low32_only", 5, 0x1234567890ABCDEF, range(8))
    add("key_zero", 0, 0, range(8))
    add("key_max32", 123, 0xFFFFFFFF, range(8))
    rng = np.random.RandomState(1234)
    for k in range(6):
        start = int(rng.randint(0, 2 ** 62, dtype=np.int64))
        key = int(rng.randint(0, 2 ** 32, dtype=np.int64))
        add("random_%d" % k, start, key, [int(x) for x in rng.randint(0, 2 ** 31, size=8)])
    path = os.path.join(HERE, "philox2x32_10.json")
    with open(path, "w") as f:
        json.dump({"generator": "reference random123() via oracle/_ref/libr123_ref.so",
                   "algorithm": "philox2x32, 10 rounds; ctr = start + index; key = low 32 bits",
                   "cases": cases}, f, indent=1)
    print("wrote", path, len(cases), "cases")


if __name__ == "__main__":
    main()
