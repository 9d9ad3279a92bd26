"""A lossless JPEG (ITU T.81 process 14) decoder for the streams DNG tiles
hold, written from the standard: SOF3 frame header, one DHT table, SOS with
predictor 1 and no point transform, 16-bit precision, components
interleaved one sample each. Test helper: the independent reader that the
encoder's round trips are checked with."""


def _be16(data, at):
    return (data[at] << 8) | data[at + 1]


def decode(data: bytes):
    """(samples as a flat list in CFA order, width in sensels, height,
    components)."""
    assert data[0:2] == b"\xff\xd8", "SOI"
    at = 2
    width = height = components = None
    table = {}
    while True:
        assert data[at] == 0xFF
        marker = data[at + 1]
        length = _be16(data, at + 2)
        body = data[at + 4:at + 2 + length]
        at += 2 + length
        if marker == 0xC3:  # SOF3
            assert body[0] == 16
            height, width, components = _be16(body, 1), _be16(body, 3), body[5]
        elif marker == 0xC4:  # DHT: class/id, BITS[16], HUFFVAL
            assert body[0] == 0
            counts = body[1:17]
            values = body[17:17 + sum(counts)]
            code, k = 0, 0
            for bits, n in enumerate(counts, start=1):
                for _ in range(n):
                    table[(bits, code)] = values[k]
                    code += 1
                    k += 1
                code <<= 1
        elif marker == 0xDA:  # SOS
            assert body[0] == components
            predictor, point = body[1 + 2 * components], body[3 + 2 * components]
            assert predictor == 1 and point == 0
            break
        else:
            raise AssertionError(f"marker {marker:#x}")
    # Entropy-coded segment up to EOI, with the stuffed zero after 0xFF removed.
    seg = bytearray()
    while not (data[at] == 0xFF and data[at + 1] == 0xD9):
        seg.append(data[at])
        if data[at] == 0xFF:
            assert data[at + 1] == 0
            at += 2
        else:
            at += 1
    assert at + 2 == len(data), "EOI ends the stream"
    bitstring = "".join(f"{b:08b}" for b in seg)
    pos = 0

    def take(n):
        nonlocal pos
        v = int(bitstring[pos:pos + n], 2) if n else 0
        pos += n
        return v

    full = width * components
    out = [0] * (full * height)
    for i in range(full * height):
        length, code = 0, 0
        while True:
            code = (code << 1) | take(1)
            length += 1
            if (length, code) in table:
                t = table[(length, code)]
                break
            assert length < 16, "no code"
        if t == 0:
            diff = 0
        elif t == 16:
            diff = 32768
        else:
            v = take(t)
            diff = v if v >= 1 << (t - 1) else v - (1 << t) + 1  # T.81 F.2.2.1 EXTEND
        row, col = divmod(i, full)
        if col >= components:
            pred = out[i - components]
        elif row > 0:
            pred = out[i - full]
        else:
            pred = 1 << 15
        out[i] = (pred + diff) & 0xFFFF
    # Whatever is left is padding of one bits to the byte boundary.
    rest = bitstring[pos:]
    assert len(rest) < 8 and set(rest) <= {"1"}, "padding"
    return out, full, height, components
