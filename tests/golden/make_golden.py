"""Writes the committed golden fixtures under tests/golden/.

Run once in a container that has the reference checkout (read as data only:
JSON, CSV and Arrow IPC files its own tests hold, plus constants quoted from
its test sources).  Nothing here imports or runs reference code; independent
values come from the Python ``xxhash`` package and from pure-Python
restatements of published algorithms (XXH64, the Parquet split-block Bloom
filter specification).

    python tests/golden/make_golden.py /root/reference
"""
from __future__ import annotations

import hashlib
import json
import os
import random
import shutil
import struct
import sys

import xxhash

HERE = os.path.dirname(os.path.abspath(__file__))


def xxh64(b: bytes) -> int:
    return xxhash.xxh64_intdigest(b, seed=0)


# ---- Parquet SBBF from the specification (independent of the oracle) ------
# parquet-format BloomFilter.md: block = ((h >> 32) * num_blocks) >> 32,
# key = low 32 bits of h, bit of word i = (key * SALT[i]) >> 27.
SALT = [0x47B6137B, 0x44974D91, 0x8824AD5B, 0xA2B7289D,
        0x705495C7, 0x2DF1424B, 0x9EFC4947, 0x5C6BFB31]


def sbbf_insert(words, nblocks, h):
    b = ((h >> 32) * nblocks) >> 32
    key = h & 0xFFFFFFFF
    for i in range(8):
        words[b * 8 + i] |= 1 << (((key * SALT[i]) & 0xFFFFFFFF) >> 27)


def sbbf_bytes(words):
    return b"".join(struct.pack("<I", w) for w in words)


def main(ref: str):
    out = {}
    # -- routing hash vectors quoted from crates/krishiv-kernels/tests/routing.rs:30-145
    out["routing"] = {
        "source": "crates/krishiv-kernels/tests/routing.rs:30-145",
        "xxh64": {"": 0xEF46DB3751D8E999, "a": 0xD24EC4F1A98C6E5B, "abc": 0x44BC2CF5AD770999},
        "int64": {"values": [0, 1, -1, 1 << 40],
                  "hashes": [3803688792395291579, 11468921228449061269,
                             9642548396912002761, 11618905265722610771]},
        "null_hash": 0xA5A55A5AC3C33C3C,
        "int32": {"values": [7, -7], "hashes": [13243476881813409038, 7562510126331999253]},
        "float64": {"values": ["0.0", "-0.0", "nan", "-nan", "1.5"],
                    "hashes": [3803688792395291579, 3803688792395291579,
                               16838308782748609196, 16838308782748609196,
                               5329932555030153977]},
        "decimal128": {"values": [12345, -1], "precision": 20, "scale": 2,
                       "hashes": [13438840123374532013, 12042982046572201028]},
        "string": {"value": "krishiv key group", "hash": 807839895623716763},
        "two_columns": {"int64": [1, 2], "string": ["x", "y"],
                        "hashes": [17888868210822470624, 17970421014506146441]},
        "fold_init_0": 79260605269340283,
    }
    # -- independent XXH64 values over every fixed width (python xxhash)
    rng = random.Random(20261203)
    fixed = []
    for _ in range(200):
        v = rng.getrandbits(128)
        b = v.to_bytes(16, "little")
        fixed.append({"v": str(v), "w1": xxh64(b[:1]), "w2": xxh64(b[:2]), "w4": xxh64(b[:4]),
                      "w8": xxh64(b[:8]), "w16": xxh64(b)})
    out["xxh64_fixed"] = fixed
    lens = [0, 1, 3, 4, 7, 8, 15, 16, 31, 32, 33, 63, 64, 100, 257]
    gen = []
    for L in lens:
        b = bytes(rng.getrandbits(8) for _ in range(L))
        gen.append({"hex": b.hex(), "seed0": xxh64(b),
                    "seed42": xxhash.xxh64_intdigest(b, seed=42)})
    out["xxh64_bytes"] = gen

    # -- row-layout hash known answer (crates/krishiv-exec/src/join/keys.rs:348)
    out["row_hash"] = {"source": "crates/krishiv-exec/src/join/keys.rs:326-348",
                       "keys": "float64 0.0, int32 NULL (null-safe)",
                       "h": 4733552696251797926}

    # -- SBBF: the reference's arrow-rs agreement case (bloom.rs:329-347),
    #    regenerated from the Parquet spec
    nb = 1024 // 32
    words = [0] * (nb * 8)
    for i in range(500):
        v = i * 7919 - 3
        sbbf_insert(words, nb, xxh64(struct.pack("<q", v)))
    out["sbbf_1024"] = {"source": "crates/krishiv-io/src/runtime_filter/bloom.rs:329-347",
                        "values": "i*7919-3 for i<500, XXH64 of 8 LE bytes",
                        "bytes_hex": sbbf_bytes(words).hex()}
    # config-3-sized filter: 750,000 keys at fpp 0.03 -> 723,968 bytes; keep a digest
    nbytes = 723968
    nb = nbytes // 32
    words = [0] * (nb * 8)
    for k in range(750000):
        sbbf_insert(words, nb, xxh64(struct.pack("<q", k * 20 + 3)))
    out["sbbf_config3"] = {"items": 750000, "fpp": 0.03, "bytes": nbytes,
                           "keys": "k*20+3 for k<750000",
                           "sha256": hashlib.sha256(sbbf_bytes(words)).hexdigest()}

    with open(os.path.join(HERE, "vectors.json"), "w") as f:
        json.dump(out, f, indent=1)

    # -- Spark 4.2.0 decimal sum/avg goldens (data file of the reference's tests)
    shutil.copyfile(os.path.join(ref, "tests/fixtures/spark-4.2.0/decimal-sum-avg.json"),
                    os.path.join(HERE, "spark_decimal_sum_avg.json"))
    # -- Spark 4.2.0 hash()/xxhash64() of single int and bigint values (the int
    # and bigint cases of the reference's data file): Murmur3 x86_32 and
    # XXH64, both seed 42
    with open(os.path.join(ref, "tests/fixtures/spark-4.2.0/hash.json")) as f:
        hj = json.load(f)
    cases = [{"type": c["category"], "value": int(c["expr"].rstrip("L")), "hash": c["hash"],
              "xxhash64": c["xxhash64"]}
             for c in hj["cases"] if c["category"] in ("int", "bigint") and not c["conf"]]
    with open(os.path.join(HERE, "spark_hash_int_bigint.json"), "w") as f:
        json.dump({"source": "tests/fixtures/spark-4.2.0/hash.json (int and bigint cases)",
                   "spark_version": hj["spark_version"], "cases": cases}, f, indent=1)
    # -- TPC-H Q1 aggregate input (Spark 4.2.0, SF 0.01) and Spark's answer
    tp = os.path.join(ref, "crates/krishiv-exec/tests/fixtures/tpch-sf0.01")
    for name in ("q1_input.arrow", "q1_expected.csv"):
        shutil.copyfile(os.path.join(tp, name), os.path.join(HERE, "tpch_" + name))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
