"""Helpers shared by the golden-vector tests."""
import hashlib


def digest(keys, values):
    """SHA-256 of a map's little-endian key words followed by its binary64 values."""
    h = hashlib.sha256()
    h.update(keys.astype("<u8").tobytes())
    h.update(values.astype("<f8").tobytes())
    return h.hexdigest()
