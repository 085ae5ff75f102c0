"""An independent decoder of the codec's compressed stream, written from the
stream's layout (csrc/pmsz_dqstream.h) and not from the oracle's code: struct
parsing, a dictionary of canonical codes, bits read from Python integers, and the
inverse Lorenzo transform as numpy cumulative sums (mod 2^32)."""
import struct

import numpy as np

SYMBOLS, ESCAPE, RADIUS, CHUNK = 1025, 1024, 512, 2048


def split(n, parts, coord):
    base, rem = divmod(n, parts)
    return base + (1 if coord < rem else 0), coord * base + min(coord, rem)


def canonical(lengths):
    """{(length, code): symbol}: codes of one length consecutive in symbol order,
    shorter codes first (DEFLATE)."""
    out, code, prev = {}, 0, 0
    for ln, sym in sorted((ln, s) for s, ln in enumerate(lengths) if ln):
        code <<= ln - prev
        out[(ln, code)] = sym
        code += 1
        prev = ln
    return out


def decode(stream):
    b = bytes(stream)
    magic, dt, nx, ny, nz, s0, s1, s2, nb, eb, chunk = struct.unpack_from("<IIqqqiiiIdI", b, 0)
    assert magic == 0x31514450 and chunk == CHUNK
    T = np.float32 if dt == 0 else np.float64
    recs = [struct.unpack_from("<QQQQQII", b, 64 + 64 * r) for r in range(nb)]
    o_len = 64 + 64 * nb
    o_cwo = (o_len + nb * SYMBOLS + 7) // 8 * 8
    chunks = sum(r[6] for r in recs)
    nesc = sum(r[1] for r in recs)
    nverb = sum(r[2] for r in recs)
    o_ceo = o_cwo + 8 * (chunks + 1)
    o_bits = o_ceo + 8 * chunks
    cwo = struct.unpack_from(f"<{chunks + 1}Q", b, o_cwo)
    ceo = struct.unpack_from(f"<{chunks}Q", b, o_ceo)
    o_esc = o_bits + 4 * cwo[-1]
    o_vidx = o_esc + 4 * nesc
    o_vval = (o_vidx + 4 * nverb + 7) // 8 * 8
    escs = struct.unpack_from(f"<{nesc}I", b, o_esc)
    vidx = struct.unpack_from(f"<{nverb}I", b, o_vidx)
    vval = np.frombuffer(b, T, nverb, o_vval)
    g0 = np.empty((nz, ny, nx), T)
    vo = 0
    for r in range(nb):
        cx, cy, cz = r // (s1 * s2), (r // s2) % s1, r % s2
        (lx, ox), (ly, oy), (lz, oz) = split(nx, s0, cx), split(ny, s1, cy), split(nz, s2, cz)
        vox, _, nverb_b, _, _, first, nch = recs[r]
        book = canonical(b[o_len + r * SYMBOLS: o_len + (r + 1) * SYMBOLS])
        d = np.empty(vox, np.uint64)
        for c in range(nch):
            w0, w1 = cwo[first + c], cwo[first + c + 1]
            bits = int.from_bytes(b[o_bits + 4 * w0: o_bits + 4 * w1], "little")
            pos, e = 0, ceo[first + c]
            for u in range(c * CHUNK, min((c + 1) * CHUNK, vox)):
                code, ln = 0, 0
                while (ln, code) not in book:
                    code = (code << 1) | ((bits >> pos) & 1)
                    pos += 1
                    ln += 1
                    assert ln <= 24, "no code"
                sym = book[(ln, code)]
                if sym == ESCAPE:
                    d[u] = escs[e]
                    e += 1
                else:
                    d[u] = (sym - RADIUS) & 0xffffffff
        k = d.reshape(lz, ly, lx)
        for axis in (2, 1, 0):
            k = np.cumsum(k, axis=axis, dtype=np.uint64) & 0xffffffff
        kk = k.astype(np.uint32).view(np.int32).astype(np.float64)
        blk = (kk * (2.0 * eb)).astype(T)
        for _ in range(nverb_b):
            li = vidx[vo]
            blk.reshape(-1)[li] = vval[vo]
            vo += 1
        g0[oz:oz + lz, oy:oy + ly, ox:ox + lx] = blk
    return g0, (nesc, nverb, cwo[-1])
