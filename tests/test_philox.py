"""Pins the Philox4x32-10 implementation (harmonize_amd/specs/philox.h, shared by
the runtime's specifications and the oracle) against rocRAND's independent
implementation, including the three Random123 known-answer vectors."""

import json
from pathlib import Path

import oracle_lib as O

VECTORS = json.loads((Path(__file__).parent / "golden" / "philox4x32_10.json").read_text())["vectors"]


def test_random123_known_answers():
    assert O.philox([0, 0, 0, 0], [0, 0]) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert O.philox([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert O.philox([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0]) == [
        0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def test_matches_rocrand_vectors():
    assert len(VECTORS) == 64
    for v in VECTORS:
        assert O.philox(v["counter"], v["key"]) == v["out"], v
