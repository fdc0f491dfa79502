"""Writes rng_lowbias32.json: hash values and randomFloat sequences of the
reference RNG (PathTrace.metal:264-276), computed with Python integers from
Wellons' published lowbias32 (hash-prospector: x ^= x >> 16; x *= 0x7feb352d;
x ^= x >> 15; x *= 0x846ca68b; x ^= x >> 16) and the per-pixel seeding
(PathTrace.metal:5212-5214)."""
import json
import os

M = 0xFFFFFFFF


def h(x):
    x ^= x >> 16
    x = (x * 0x7FEB352D) & M
    x ^= x >> 15
    x = (x * 0x846CA68B) & M
    x ^= x >> 16
    return x


def main():
    hashes = {str(v): h(v) for v in [0, 1, 2, 3, 255, 0x1234ABCD, 0xDEADBEEF, 0xFFFFFFFF, 57599, 2073599]}
    seqs = []
    for frame_seed, px, py, width, sample in [(0, 0, 0, 320, 0), (0, 17, 9, 320, 0), (0x1234ABCD, 5, 3, 48, 3),
                                              (0, 1919, 1079, 1920, 63)]:
        st = frame_seed ^ h(px + py * width) ^ h(sample + 1)
        seed = st
        vals = []
        for _ in range(8):
            st = h(st)
            vals.append((st & 0xFFFFFF) / float(1 << 24))
        seqs.append({"frame_seed": frame_seed, "px": px, "py": py, "width": width, "sample": sample, "seed": seed,
                     "floats": vals, "state_after": st})
    out = {"hash": hashes, "sequences": seqs}
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "rng_lowbias32.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
