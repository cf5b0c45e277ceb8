"""CPU oracle (pure Python) of the reference's Cuper packed matrix format.

TEST INFRASTRUCTURE ONLY (imported by tests/, never by the product).
Restates host/cuper_control_matrix.hpp:

  pe_for_row      cuper_physical_pe_for_row (:48-54)
  restore_row     the kernel's inverse, restore_global_row
                  (kernels/cuper_pcg_control_kernel.cpp:140-147)
  pack            build_cuper_control_matrix (:136-207) with
                  append_reordered_cuper_elements (:82-127) and
                  pack_cuper_element (:73-80)
  unpack          the library's import: every non-empty word back to
                  (row, column, FP32 value)

One difference is unavoidable: the reference sorts each PE list with
std::sort (not stable), and the order it leaves entries of equal column in
is libstdc++'s introsort's.  This restatement sorts stably, so it equals
the reference word for word only where no PE list holds two entries of one
column within a batch; tests/test_cuper_format.py checks it against the
reference's own packer (oracle/_ref, built from host/cuper_control_matrix.hpp)
on such matrices, and checks the library against the reference's packer on
all matrices.  The format invariants (window of 10 slots per row group,
padding, cumulative batch table) are checked on every packing.
"""
from __future__ import annotations

import struct

SLICE = 8192
CHANNELS = 16
LANES = 8
PES = CHANNELS * LANES
WINDOW = 10
EMPTY = 0x3FFFF << 32


def pe_for_row(row):
    packet = row // 2
    return ((packet % 8) * 2 + (packet // 8) % 2) * 8 + (packet // 16) % 8


def restore_row(channel, lane, code):
    packet = (code >> 1) * PES + channel // 2 + (channel & 1) * 8 + lane * 16
    return packet * 2 + (code & 1)


def f32_bits(v):
    return struct.unpack("<I", struct.pack("<f", v))[0]


def bits_f32(b):
    return struct.unpack("<f", struct.pack("<I", b))[0]


def word(col, code, value):
    return ((col & 0x3FFF) << 50) | ((code & 0x3FFFF) << 32) | f32_bits(value)


def pack(n, row_ptr, col_idx, values):
    """(ptr, data) with data a list of 16 channel lists."""
    batches = (n + SLICE - 1) // SLICE
    bucket = [[[] for _ in range(PES)] for _ in range(batches)]
    for r in range(n):
        pe = pe_for_row(r)
        for o in range(row_ptr[r], row_ptr[r + 1]):
            c = int(col_idx[o])
            bucket[c // SLICE][pe].append((c, r, float(values[o])))
    lists = [[] for _ in range(PES)]
    ptr = [0]
    for b in range(batches):
        for pe in range(PES):
            ent = sorted(bucket[b][pe], key=lambda e: e[0])
            sched = []
            last = {}
            for c, r, v in ent:
                g = r // (2 * PES)
                slot = last.get(g, -WINDOW) + WINDOW
                while slot < len(sched) and sched[slot] is not None:
                    slot += 1
                if slot >= len(sched):
                    sched.extend([None] * (slot + 1 - len(sched)))
                sched[slot] = (c - b * SLICE, g * 2 + (r & 1), v)
                last[g] = slot
            lists[pe].extend(sched)
        longest = max(len(x) for x in lists)
        for x in lists:
            x.extend([None] * (longest - len(x)))
        ptr.append(longest)
    length = ptr[-1]
    size = (8 * length + 511) // 512 * 512
    data = [[0] * size for _ in range(CHANNELS)]
    for i in range(length):
        for ch in range(CHANNELS):
            for lane in range(LANES):
                e = lists[ch * LANES + lane][i]
                data[ch][i * 8 + lane] = EMPTY if e is None else word(*e)
    return ptr, data


def unpack(ptr, data, n):
    """CSR (row_ptr, col_idx, values) with columns ascending in each row."""
    ent = []
    for ch in range(CHANNELS):
        for lane in range(LANES):
            for b in range(len(ptr) - 1):
                for i in range(ptr[b], ptr[b + 1]):
                    w = int(data[ch][i * 8 + lane])
                    code = (w >> 32) & 0x3FFFF
                    if code & (1 << 17):
                        continue
                    r = restore_row(ch, lane, code)
                    c = b * SLICE + (w >> 50)
                    if r >= n or c >= n:
                        raise ValueError("packed entry outside the matrix")
                    ent.append((r, c, bits_f32(w & 0xFFFFFFFF)))
    ent.sort(key=lambda e: (e[0], e[1]))
    row_ptr = [0] * (n + 1)
    for r, _, _ in ent:
        row_ptr[r + 1] += 1
    for r in range(n):
        row_ptr[r + 1] += row_ptr[r]
    return row_ptr, [e[1] for e in ent], [e[2] for e in ent]


def check_invariants(ptr, data, n):
    """The format's rules on any packing: window >= 10 between entries of a
    row group in one PE list and batch, empty slots as 0x3FFFF << 32, the
    words past matrix_len zero, ptr cumulative."""
    assert ptr[0] == 0 and all(a <= b for a, b in zip(ptr, ptr[1:]))
    length = ptr[-1]
    for ch in range(CHANNELS):
        assert all(int(w) == 0 for w in data[ch][8 * length:])
        for lane in range(LANES):
            for b in range(len(ptr) - 1):
                last = {}
                for i in range(ptr[b], ptr[b + 1]):
                    w = int(data[ch][i * 8 + lane])
                    code = (w >> 32) & 0x3FFFF
                    if code & (1 << 17):
                        assert w == EMPTY
                        continue
                    g = code >> 1
                    if g in last:
                        assert i - last[g] >= WINDOW
                    last[g] = i
