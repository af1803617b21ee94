"""Write tests/golden/reference_vectors.json from the reference's own CPU code.

Runs in the build container only (needs oracle/_ref/libmans_ref.so, compiled from the
reference sources by `make -C oracle ref`).  The JSON holds data only: counts, decoder
tables, symbols, encoded units and the MulticoreDecoder output for small demo-shaped
cases (src/main.cc:55-146 builds the same objects).
"""
import ctypes as C
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import oracle_lib as ol  # noqa: E402


def main():
    R = ol.ref()
    cases = []
    for lam, L, n, S, T in [(0.1, 1024, 3000, 4, 4), (1.06, 1024, 2500, 4, 2),
                            (2.34, 1024, 1777, 8, 3), (0.26, 4096, 4000, 4, 4)]:
        nsym = 256
        counts = np.zeros(nsym, np.uint32)
        dec = np.zeros(L, np.uint32)
        en, ec, el = (np.zeros(nsym * L, np.uint32) for _ in range(3))
        R.ref_demo_tables(lam, nsym, L, 5, counts, dec, en, ec, el)
        data = np.zeros(n + 8, np.uint8)
        R.ref_gen_test_data(counts, nsym, n, L, 5, data)
        units = np.zeros(n, np.uint32)
        fb, fs = C.c_uint32(), C.c_uint32()
        nu = R.ref_encode(counts, nsym, L, data, n, units, len(units), C.byref(fb), C.byref(fs))
        units = units[:nu]
        out = np.zeros(n, np.uint8)
        R.ref_multicore(S, T, np.concatenate([units, np.zeros(S + 8, np.uint32)]), nu, dec, L,
                        fs.value, fb.value, out, n)
        cases.append({"lambda": lam, "L": L, "S": S, "T": T, "counts": counts.tolist(),
                      "dec_table": dec.tolist(), "symbols_hex": data[:n].tobytes().hex(),
                      "units": units.tolist(), "first_bit": fb.value, "first_state": fs.value,
                      "multicore_hex": out.tobytes().hex()})
    with open(os.path.join(HERE, "reference_vectors.json"), "w") as f:
        json.dump({"source": "reference ANSTableGenerator/ANSEncoder/MulticoreDecoder via oracle/_ref",
                   "cases": cases}, f)


if __name__ == "__main__":
    main()
