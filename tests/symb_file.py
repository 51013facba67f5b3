"""Reader/writer of the reference's .symb layout for the tests
(persistent_solution.cu:31-199): 16 size_t counts, then int32 / float32 arrays."""
import struct

import numpy as np

ORDER = [("rows", "i"), ("cols", "i"), ("indptr", "i"), ("ordered_vertices", "i"), ("terminals_offsets", "i"),
         ("nonterminals_offsets", "i"), ("rates", "f"), ("term_indptr", "i"), ("term_indices", "i"),
         ("term_data", "f"), ("nonterm_indptr", "i"), ("nonterm_indices", "i"), ("nonterm_data", "f"),
         ("inv_indptr", "i"), ("inv_indices", "i"), ("inv_data", "f")]


def write(path, d):
    arrays = [np.asarray(d.get(k, []), dtype=np.int32 if t == "i" else np.float32) for k, t in ORDER]
    with open(path, "wb") as f:
        f.write(struct.pack("<16Q", *[a.size for a in arrays]))
        for a in arrays:
            f.write(a.tobytes())


def read(path):
    raw = open(path, "rb").read()
    sizes = struct.unpack_from("<16Q", raw, 0)
    out, off = {}, 128
    for (k, t), n in zip(ORDER, sizes):
        dt = np.int32 if t == "i" else np.float32
        out[k] = np.frombuffer(raw, dtype=dt, count=n, offset=off)
        off += 4 * n
    assert off == len(raw), "length does not match the counts"
    return out
