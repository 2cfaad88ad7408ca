"""Determinant codec.

Reference layout (Core/WaveFunctionArithmetic.cpp:25-121, Core/OperatorWaveFunction.cpp:10-12):
mode i is bit 128>>(i&7) of byte i>>3; a determinant string is a '0'/'1' per
mode, spaces ignored. libqs layout: NW uint64 words (NW = 1 for NF <= 64, 2 for
NF <= 128), word 0 most significant, mode i at bit 64*NW-1-i — the reference
bytes read big-endian, so numeric order == memcmp order of the bytes ==
lexicographic order of the strings.
"""
from __future__ import annotations

import numpy as np


def nwords(nf: int) -> int:
    if nf < 1 or nf > 128:
        raise ValueError("NF must be in 1..128")
    return 1 if nf <= 64 else 2


def nbytes(nf: int) -> int:
    return (nf + 7) // 8


def string_to_words(s: str, nf: int) -> tuple:
    """'0101 10' -> tuple of NW ints (StringToDet, Core/WaveFunctionArithmetic.cpp:25)."""
    bits = [c for c in s if c in "01"]
    if any(c not in "01 " for c in s):
        raise ValueError(f"determinant string may hold only 0, 1 and spaces: {s!r}")
    if len(bits) != nf:
        raise ValueError(f"determinant string has {len(bits)} modes, expected {nf}")
    nw = nwords(nf)
    v = 0
    for c in bits:
        v = (v << 1) | (c == "1")
    v <<= 64 * nw - nf
    return tuple((v >> (64 * (nw - 1 - w))) & 0xFFFFFFFFFFFFFFFF for w in range(nw))


def words_to_string(words, nf: int) -> str:
    """(DetToString, Core/WaveFunctionArithmetic.cpp:96)."""
    nw = nwords(nf)
    v = 0
    for w in range(nw):
        v = (v << 64) | int(words[w])
    v >>= 64 * nw - nf
    return format(v, f"0{nf}b")


def modes_to_words(modes, nf: int) -> tuple:
    nw = nwords(nf)
    v = 0
    for i in modes:
        v |= 1 << (64 * nw - 1 - i)
    return tuple((v >> (64 * (nw - 1 - w))) & 0xFFFFFFFFFFFFFFFF for w in range(nw))


def words_to_int(words) -> int:
    v = 0
    for w in words:
        v = (v << 64) | int(w)
    return v


def bytes_to_words(b: np.ndarray, nf: int) -> np.ndarray:
    """(n, nbytes) uint8 reference bytes -> (n, NW) uint64 words."""
    nb, nw = nbytes(nf), nwords(nf)
    b = np.asarray(b, dtype=np.uint8).reshape(-1, nb)
    pad = np.zeros((b.shape[0], 8 * nw), dtype=np.uint8)
    pad[:, :nb] = b
    out = np.zeros((b.shape[0], nw), dtype=np.uint64)
    for w in range(nw):
        chunk = pad[:, 8 * w:8 * w + 8].astype(np.uint64)
        acc = np.zeros(b.shape[0], dtype=np.uint64)
        for k in range(8):
            acc = (acc << np.uint64(8)) | chunk[:, k]
        out[:, w] = acc
    return out


def words_to_bytes(words: np.ndarray, nf: int) -> np.ndarray:
    nb, nw = nbytes(nf), nwords(nf)
    words = np.asarray(words, dtype=np.uint64).reshape(-1, nw)
    out = np.zeros((words.shape[0], 8 * nw), dtype=np.uint8)
    for w in range(nw):
        for k in range(8):
            out[:, 8 * w + k] = ((words[:, w] >> np.uint64(56 - 8 * k)) & np.uint64(0xFF)).astype(np.uint8)
    return out[:, :nb].copy()


def mask_bytes(mask: str, nf: int) -> np.ndarray:
    """Restriction mask string -> reference bytes (spaces ignored, TableToArray.cpp:3005-3068)."""
    w = string_to_words(mask, nf)
    return words_to_bytes(np.array([w], dtype=np.uint64), nf)[0]
