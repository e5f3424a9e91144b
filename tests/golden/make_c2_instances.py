"""Draw the 64 OneChQS Anderson instances of the batched configuration:
splitmix64 with seed 20261203, u = (x >> 11) * 2^-53, draws per instance in the
order (U, ed, Gamma, Ncut): U = 0.1 + 0.9 u1, ed = -U u2, Gamma = 0.01 + 0.09 u3,
Ncut = 800 + floor(1201 u4); Lambda = 2.5, 45 sites.  Writes c2_instances.json."""
import json
import os

MASK = (1 << 64) - 1


def splitmix64(state):
    while True:
        state = (state + 0x9E3779B97F4A7C15) & MASK
        z = state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
        yield z ^ (z >> 31)


def draw(n=64, seed=20261203):
    g = splitmix64(seed)
    u = lambda: (next(g) >> 11) * 2.0 ** -53
    out = []
    for _ in range(n):
        U = 0.1 + 0.9 * u()
        ed = -U * u()
        G = 0.01 + 0.09 * u()
        nc = 800 + int(1201 * u())
        out.append(dict(U=U, ed=ed, Gamma=G, ncut=nc, Lambda=2.5))
    return out


if __name__ == "__main__":
    here = os.path.dirname(os.path.abspath(__file__))
    with open(os.path.join(here, "c2_instances.json"), "w") as f:
        json.dump({"seed": 20261203, "nsites": 45, "instances": draw()}, f, indent=1)
