"""Writes tests/golden/generators.json: R-MAT edge heads and an order-free
digest, computed by an independent pure-Python SplitMix64 / R-MAT
implementation (not the oracle), so the C and HIP generators are pinned to the
published SplitMix64 stream and the documented quadrant rule."""
import json
import os

MASK = (1 << 64) - 1
GOLD = 0x9E3779B97F4A7C15


def fmix(z):
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
    return z ^ (z >> 31)


def stream_key(seed, stream):
    return fmix(fmix(seed) ^ ((stream * 0xD6E8FEB86659FD93) & MASK))


def rand64(key, i):
    return fmix((key + (i + 1) * GOLD) & MASK)


def rmat(key, scale, n, a=0.57, b=0.19, c=0.19):
    s = 4294967296.0
    ta, tab, tabc = int(a * s), int((a + b) * s), int((a + b + c) * s)
    out = []
    for e in range(n):
        u = v = 0
        for lvl in range(scale):
            r = rand64(key, e * scale + lvl) >> 32
            bit = 1 << (scale - 1 - lvl)
            if r < ta:
                pass
            elif r < tab:
                v |= bit
            elif r < tabc:
                u |= bit
            else:
                u |= bit
                v |= bit
        out.append((u, v))
    return out


def main():
    cases = []
    for seed, scale, n in [(1, 14, 4000), (2, 10, 2000), (3, 20, 1000)]:
        e = rmat(stream_key(seed, 1), scale, n)
        dig = 0
        for u, v in e:
            dig ^= (((u << 32) | v) * GOLD) & MASK
        cases.append({"seed": seed, "scale": scale, "n": n,
                      "src_head": [u for u, _ in e[:8]], "dst_head": [v for _, v in e[:8]],
                      "xor_digest": dig})
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "generators.json")
    with open(path, "w") as f:
        json.dump({"rmat": cases}, f, indent=1)


if __name__ == "__main__":
    main()
