"""d3-force's default random source (src/lcg.js:1-9):
``s = (1664525 s + 1013904223) mod 2^32; return s / 2^32``, starting at s = 1.
The product stays below 2^53, so integer arithmetic gives d3's doubles."""


def lcg(seed=1):
    state = [seed]

    def random():
        state[0] = (1664525 * state[0] + 1013904223) % 4294967296
        return state[0] / 4294967296

    random.state = state  # the generator's position, for saving and restoring it
    return random
