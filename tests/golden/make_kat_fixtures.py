"""Extract the reference's known-answer vectors into JSON fixtures.

Reads the via-rs test data (plain constant arrays) and the byte-stream /
sampler constants inlined in its unit tests, and writes them as data under
tests/golden/. Only data is copied: names, shapes and numbers.

    python tests/golden/make_kat_fixtures.py [/path/to/via-rs]

Sources:
  crates/via-primitives/tests/data/layer3_kats.rs   (switching)
  crates/via-primitives/tests/data/layer4_kats.rs   (gates)
  crates/via-primitives/tests/data/layer5_kats.rs   (cascade)
  crates/via-integration/tests/data/layer6_kats.rs  (protocol composites)
  crates/via-primitives/src/sampling/prg.rs          (PRG byte stream, uniform_below)
  crates/via-primitives/src/sampling/ternary.rs      (ternary vectors)
"""
import json
import os
import re
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
OUT = os.path.dirname(os.path.abspath(__file__))

ARRAY = re.compile(
    r"const\s+([A-Z0-9_]+)\s*:\s*\[\s*(u64|u8|i8)\s*;\s*[0-9_]+\s*\]\s*=\s*\[(.*?)\];",
    re.S,
)


def parse_int(tok):
    tok = tok.strip().replace("_", "")
    if not tok:
        return None
    return int(tok, 16) if tok.lower().startswith(("0x", "-0x")) else int(tok)


def arrays(path, names=None):
    src = open(path).read()
    out = {}
    for name, _ty, body in ARRAY.findall(src):
        if names is not None and name not in names:
            continue
        body = re.sub(r"//[^\n]*", "", body)
        vals = [parse_int(t) for t in body.split(",")]
        out[name] = [v for v in vals if v is not None]
    return out


def main():
    fx = {}
    for layer, rel in [
        ("layer3", "crates/via-primitives/tests/data/layer3_kats.rs"),
        ("layer4", "crates/via-primitives/tests/data/layer4_kats.rs"),
        ("layer5", "crates/via-primitives/tests/data/layer5_kats.rs"),
        ("layer6", "crates/via-integration/tests/data/layer6_kats.rs"),
    ]:
        fx[layer] = arrays(os.path.join(REF, rel))
    fx["prg"] = arrays(
        os.path.join(REF, "crates/via-primitives/src/sampling/prg.rs"),
        {
            "EMPTY_SEED_FIRST_32",
            "TEST_SEED_FIRST_32",
            "TEST_SEED_BOUNDARY_134_138",
            "TEST_SEED_RANDBELOW_3_FIRST_8",
            "TEST_SEED_RANDBELOW_256_FIRST_8",
            "TEST_SEED_RANDBELOW_Q3_FIRST_4",
            "TEST_SEED_RANDBELOW_2P32_FIRST_4",
            "TEST_SEED_RANDBELOW_2P32_PLUS_1_FIRST_4",
        },
    )
    fx["ternary"] = arrays(
        os.path.join(REF, "crates/via-primitives/src/sampling/ternary.rs"),
        {"TEST_SEED_TP_N16", "EMPTY_SEED_TP_N16"},
    )
    for k, v in fx.items():
        assert v, f"no arrays found for {k}"
    path = os.path.join(OUT, "reference_kats.json")
    with open(path, "w") as f:
        json.dump(fx, f, indent=0, sort_keys=True)
    print(path, {k: len(v) for k, v in fx.items()})


if __name__ == "__main__":
    main()
