"""Pure-Python Philox-4x32-10 (Salmon, Moraes, Dror, Shaw, SC 2011), used by the
tests to pin the device generator of xmatch_synth.h."""
M0, M1 = 0xD2511F53, 0xCD9E8D57
W0, W1 = 0x9E3779B9, 0xBB67AE85
MASK = 0xFFFFFFFF


def philox4x32_10(ctr, key):
    c = list(ctr)
    k0, k1 = key
    for _ in range(10):
        p0 = M0 * c[0]
        p1 = M1 * c[2]
        c = [((p1 >> 32) ^ c[1] ^ k0) & MASK, p1 & MASK, ((p0 >> 32) ^ c[3] ^ k1) & MASK, p0 & MASK]
        k0 = (k0 + W0) & MASK
        k1 = (k1 + W1) & MASK
    return c


def uniforms(seed, stream, row, draw):
    """The two 53-bit uniforms the device draws for (seed, stream, row, draw)."""
    ctr = [row & MASK, (row >> 32) & MASK, stream & MASK, draw & MASK]
    key = [seed & MASK, ((seed >> 32) ^ (stream >> 32)) & MASK]
    o = philox4x32_10(ctr, key)
    x = (o[0] << 32) | o[1]
    y = (o[2] << 32) | o[3]
    return (x >> 11) * 2.0 ** -53, (y >> 11) * 2.0 ** -53
