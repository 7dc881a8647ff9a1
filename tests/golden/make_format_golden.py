"""Write tests/golden/result_formats.json: the text line Result<int,
float>::str writes (include/fiuncho/utils/Result.h:37-45) and the bytes
Result::serialize writes (Result.h:47-55) for given results, computed by the
reference compiled from its own sources (oracle/_ref/libfiuncho_ref.so).

Results: the reference's own top-90 lists of tests/data at orders 2 to 9
(testdata_search.json), and single values chosen for the formatting corner
cases (zero, negative zero, a tiny negative FP32 MI, exponents, rounding at
the sixth significant digit, large indices and orders).

Usage:  python tests/golden/make_format_golden.py
"""
import json
import os
import struct
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402


def main():
    ref = oracle.ref("avx512")
    if ref is None:
        sys.exit("oracle/_ref is not built (make -C oracle ref)")
    with open(os.path.join(HERE, "testdata_search.json")) as f:
        search = json.load(f)
    cases = []
    for k, g in sorted(search["top"].items(), key=lambda t: int(t[0])):
        for c, v in zip(g["combos"], g["mi"]):
            cases.append((c, v))
    specials = [0.0, -0.0, -1.1920929e-07, 1e-05, 9.9999997e-06, 0.1234565, 0.12345651, 1.0, 0.6931472,
                123456.7, 1234567.0, 3.4028235e38, 1e-30, 1.4e-45, 0.022601502016186714, 0.0001]
    for i, v in enumerate(specials):
        cases.append(([i, i + 1, 10 + 7 * i], v))
    cases.append(([0, 1], 0.5))
    cases.append(([2147483000, 2147483646], 0.25))
    cases.append((list(range(0, 16, 1)), 0.75))
    out = []
    for c, v in cases:
        fv = struct.unpack("<f", struct.pack("<f", v))[0]   # the float the reference holds
        out.append({"combo": c, "mi_f32_hex": struct.pack("<f", fv).hex(),
                    "str": ref.result_str(c, fv).decode(),
                    "serialize_hex": ref.result_serialize(c, fv).hex()})
    with open(os.path.join(HERE, "result_formats.json"), "w") as f:
        json.dump({"build": "g++ -O3 -march=x86-64-v4 -mavx512vpopcntdq -std=gnu++17 -DALIGN=64",
                   "results": out}, f, indent=0)
    print(f"{len(out)} results written")


if __name__ == "__main__":
    main()
