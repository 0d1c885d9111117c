"""The C++ drop-in program compiles against include/ alone and links
against the library: every template specialisation it uses (convert and
sparse_mv for the eight formats, both value types, all four stages) is
exported (no GPU needed to link)."""
from cxx_build import build_drop_in


def test_cxx_drop_in_program_compiles_and_links(tmp_path):
    exe = build_drop_in(tmp_path)
    assert exe.exists() and exe.stat().st_size > 0
