"""Test encoder of the hybrid code, restating the crate's test helper
bits::testing::BitWriter (bits.rs:138-186)."""


class BitWriter:
    def __init__(self):
        self.bits = []

    def put(self, value, n):
        for i in range(n - 1, -1, -1):
            self.bits.append((value >> i) & 1)

    def symbol(self, cb, v):
        """Append the codeword of v under codebook cb; return its length."""
        s, e, r = cb >> 8, (cb >> 4) & 15, cb & 15
        q = v >> r
        if q <= s:
            self.put(0, q)
            self.put(1, 1)
            self.put(v & ((1 << r) - 1), r)
            return q + 1 + r
        rest = v - ((s + 1) << r)
        m = rest + (1 << e)
        bits = m.bit_length()
        zeros = s + 1 + (bits - 1 - e)
        self.put(0, zeros)
        self.put(m, bits)
        return zeros + bits

    def finish(self):
        out = bytearray((len(self.bits) + 7) // 8)
        for i, b in enumerate(self.bits):
            if b:
                out[i // 8] |= 0x80 >> (i % 8)
        return bytes(out)


FIRST_DC, DC_RESET, AC_INITIAL, RN_INITIAL, LN_INITIAL = 0x2BC, 0x064, 0x031, 0x000, 0x042
DC_CB = [0x010, 0x021, 0x032, 0x033, 0x033, 0x033, 0x044, 0x044, 0x044, 0x044, 0x044, 0x044, 0x076]
LN_CB = [0x100, 0x111, 0x222, 0x222, 0x122, 0x122, 0x433, 0x433] + [0x233] * 6 + [0x033]
RN_CB = [0x200, 0x100, 0x000, 0x000, 0x211, 0x211, 0x111, 0x111, 0x011, 0x011, 0x021, 0x021,
         0x222] + [0x022] * 12 + [0x032, 0x032, 0x044]
AC_CB = ([0x000, 0x211] + [0x111] * 2 + [0x222] * 3 + [0x122] * 3 + [0x233] * 8 + [0x133] * 2 +
         [0x244] * 22 + [0x355] * 52 + [0x166])
SCAN = [0, 8, 1, 9, 16, 24, 17, 25, 2, 10, 3, 11, 18, 26, 19, 27, 32, 40, 33, 34, 41, 48, 56, 49,
        42, 35, 43, 50, 57, 58, 51, 59, 4, 12, 5, 6, 13, 20, 28, 21, 14, 7, 15, 22, 29, 36, 44,
        37, 30, 23, 31, 38, 45, 52, 60, 53, 46, 39, 47, 54, 61, 62, 55, 63]


def wrapped_zero_run_stream():
    """The stream of gpu.rs:1410-1443: +1 at scan 1, ten zeros, +1 at scan
    13, then an over-long zero run that wraps back to scan 13 where a +2
    overwrites the +1."""
    w = BitWriter()
    w.symbol(FIRST_DC, 0)
    w.symbol(LN_INITIAL, 1)
    w.symbol(AC_INITIAL, 0)
    w.put(0, 1)
    w.symbol(RN_INITIAL, 10)
    w.symbol(AC_CB[0], 0)
    w.put(0, 1)
    w.symbol(LN_CB[0], 0)
    w.put(0, 16)
    w.put(1, 1)
    w.put(0, 17)
    w.put(0b0100, 4)
    w.put(0xFFFFFFFF, 32)
    return w.finish()
