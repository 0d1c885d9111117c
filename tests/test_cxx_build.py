"""The C++ drop-in program compiles against include/ and links against the
library (no GPU needed to link)."""
from pathlib import Path


def test_cxx_drop_in_program_is_built():
    root = Path(__file__).resolve().parents[1]
    assert (root / "_build" / "api_drop_in").exists()
