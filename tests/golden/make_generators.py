"""Writes tests/golden/generators.json: R-MAT and Chung-Lu edge heads and an
order-free digest, computed by an independent pure-Python SplitMix64 / R-MAT /
Chung-Lu implementation (not the oracle), so the C and HIP generators are
pinned to the published SplitMix64 stream, the documented quadrant rule and
the inverse-CDF endpoint rule."""
import json
import math
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


def chung_lu(key, V, gamma, avg_deg, max_deg, n):
    """Endpoints by inverse-CDF lookup in the FP64 cumulative weight table
    (same operation order as the C and HIP builders: pow, running sums in
    vertex order, acc / tot * 2^62 truncated)."""
    import bisect
    ex = -1.0 / (gamma - 1.0)
    w = [math.pow(i + 1.0, ex) for i in range(V)]
    s = 0.0
    for x in w:
        s += x
    scale = avg_deg * float(V) / s
    tot = 0.0
    for i in range(V):
        w[i] *= scale
        if w[i] > max_deg:
            w[i] = max_deg
        tot += w[i]
    cum = [0] * (V + 1)
    acc = 0.0
    for i in range(V):
        acc += w[i]
        cum[i + 1] = int(acc / tot * 4611686018427387904.0)
    cum[V] = 1 << 62
    out = []
    for i in range(2 * n):
        r = rand64(key, i) >> 2
        out.append(bisect.bisect_right(cum, r) - 1)
    return out[0::2], out[1::2]


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
    cl = []
    for seed, V, gamma, avg, mx, n in [(4, 50000, 2.1, 24.0, 2000.0, 3000),
                                       (5, 1000, 2.5, 8.0, 1e9, 2000)]:
        key = stream_key(seed, 4)
        src, dst = chung_lu(key, V, gamma, avg, mx, n)
        dig = 0
        for u, v in zip(src, dst):
            dig ^= (((u << 32) | v) * GOLD) & MASK
        cl.append({"seed": seed, "V": V, "gamma": gamma, "avg_deg": avg, "max_deg": mx, "n": n,
                   "src_head": src[:8], "dst_head": dst[:8], "xor_digest": dig})
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "generators.json")
    with open(path, "w") as f:
        json.dump({"rmat": cases, "chung_lu": cl}, f, indent=1)


if __name__ == "__main__":
    main()
