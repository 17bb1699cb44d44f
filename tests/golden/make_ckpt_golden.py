"""Writes ckpt_rbm_n2_m3.json: byte-exact checkpoint fixtures of tiny RBMs in
the reference's weight-file format (cereal BinaryArchive of yannq::RBM<T>,
Serializers/SerializeRBM.hpp:21-79, SerializeEigen.hpp:10-35), built field by
field with struct.pack from cereal's binary layout: raw little-endian values,
no header; a std::unique_ptr adds a leading uint8 "valid" flag; a versioned
type adds its uint32 class version (CEREAL_CLASS_VERSION(RBM<T>, 1)) before
its fields.  Each fixture lists its fields with their offsets so that the
bytes can be checked by hand.

    python tests/golden/make_ckpt_golden.py
"""
import json
import os
import struct

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
N, M = 2, 3


def params(cplx):
    # simple dyadic values, exact in float and double: W (M x N column-major), a, b
    re = np.array([0.5, -1.25, 2.0, 0.125, -0.75, 3.5, 1.5, -2.5, 0.25, -0.5, 4.0])
    im = np.array([-0.5, 0.25, 1.0, -2.0, 0.375, 0.0, 2.25, 1.75, -1.5, 0.625, -3.0])
    return re + 1j * im if cplx else re


def encode(elem, use_bias, pointer):
    cplx = elem in ("c64", "c128")
    single = elem in ("f32", "c64")
    p = params(cplx)
    fields = []

    def put(name, fmt, *vals):
        fields.append((name, struct.pack("<" + fmt, *vals)))

    def entries(name, vals):
        f = "f" if single else "d"
        flat = []
        for v in vals:
            flat += [v.real, v.imag] if cplx else [v]
        put(name, f * len(flat), *flat)

    if pointer:
        put("unique_ptr valid (uint8)", "B", 1)
    put("class version (uint32)", "I", 1)
    put("useBias (bool)", "B", 1 if use_bias else 0)
    put("N (uint32)", "I", N)
    put("M (uint32)", "I", M)
    put("W rows = M (int32)", "i", M)
    put("W cols = N (int32)", "i", N)
    entries("W entries, column-major", p[:N * M])
    if use_bias:
        put("a rows = N (int32)", "i", N)
        put("a cols = 1 (int32)", "i", 1)
        entries("a entries", p[N * M:N * M + N])
        put("b rows = M (int32)", "i", M)
        put("b cols = 1 (int32)", "i", 1)
        entries("b entries", p[N * M + N:])
    out, layout, off = b"", [], 0
    for name, b in fields:
        layout.append([off, len(b), name])
        out += b
        off += len(b)
    dim = N * M + (N + M if use_bias else 0)
    pv = p[:N * M] if not use_bias else p
    return {"elem": elem, "use_bias": use_bias, "pointer": pointer, "n": N, "m": M,
            "params_re": [float(x) for x in np.real(pv)[:dim]],
            "params_im": [float(x) for x in np.imag(pv)[:dim]] if cplx else None,
            "hex": out.hex(), "layout": layout}


def main():
    cases = [encode(e, ub, ptr) for e in ("f32", "f64", "c64", "c128")
             for ub in (True, False) for ptr in (True, False)]
    with open(os.path.join(HERE, "ckpt_rbm_n2_m3.json"), "w") as f:
        json.dump({"source": "tests/golden/make_ckpt_golden.py", "cases": cases}, f, indent=1)


if __name__ == "__main__":
    main()
