"""CPU checks of run-time law compilation (hiprtc, no device): the built-in laws written as sources and the extra
law compile for gfx950; a law without the required members, with a syntax error or with a name that is not a C
identifier is rejected with the compiler's message (the reference's codegen raises on a form it cannot compile,
src/codegen.jl:194-212)."""
import pytest

from hfk import laws


@pytest.mark.parametrize("key", sorted(laws.SOURCES))
def test_sources_compile(key):
    src, name = laws.SOURCES[key]
    laws.check_law(src, name)


def test_missing_member_reports_compiler_log():
    with pytest.raises(ValueError, match="no member named 'tangent'"):
        laws.check_law("struct Half { __device__ void setup(const double*, double, double) {}\n"
                       "  __device__ double stress(int, int) const { return 0; } };", "Half")


def test_syntax_error_and_bad_name():
    with pytest.raises(ValueError, match="does not compile"):
        laws.check_law("struct Broken { int x }", "Broken")
    with pytest.raises(ValueError, match="C identifier"):
        laws.check_law(laws.SVK_SOURCE, "S V K")
    with pytest.raises(ValueError, match="C identifier"):
        laws.check_law(laws.SVK_SOURCE, "1SVK")
