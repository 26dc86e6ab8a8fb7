"""The generation tags of the cross-workgroup exchange (xflat_amd/csrc/
xf_step_kernel.h, word_of / tag_current), checked on the CPU by walking the
exchange sequence of the stepper.

Every step runs three exchanges, of 2 NMV (pass A), NMV (pass B) and NMV + 1
(pass C) values, NMV = 3 x the moments (MA 6, MAA 12, SA 3); exchange g uses
the rows of parity g % 2, and value v of a row lives in its own slot.  A
reader at exchange g may see a slot's previous word before the new one
lands; it must never take that word as current.  The one-word format keeps
the tag's low three bits, the two-word format the whole 32-bit tag."""
import pytest


def _writes(nmv, steps):
    """(generation, parity, value index) of every slot write, from a cleared
    state (generation 0, a step starting with pass A)."""
    out = []
    for g in range(3 * steps):
        nt = (2 * nmv, nmv, nmv + 1)[g % 3]
        out += [(g, g % 2, v) for v in range(nt)]
    return out


@pytest.mark.parametrize("nmv", [6, 12, 3])
@pytest.mark.parametrize("bits", [3, 32])
def test_no_previous_word_carries_the_current_tag(nmv, bits):
    mask = (1 << bits) - 1
    last = {}                                   # slot -> generation of its last write
    for g, p, v in _writes(nmv, 40):
        prev = last.get((p, v))
        tag = (g + 1) & mask
        seen = 0 if prev is None else (prev + 1) & mask    # a cleared slot holds 0
        assert seen != tag, (nmv, bits, g, v, prev)
        last[(p, v)] = g


def test_two_tag_bits_would_not_do():
    """With two bits a pass-A-and-C word written 4 generations before (pass C
    of the previous step on the same parity, then pass A) carries the current
    tag: the reason for three."""
    clash = []
    last = {}
    for g, p, v in _writes(6, 10):
        prev = last.get((p, v))
        if prev is not None and ((prev + 1) & 3) == ((g + 1) & 3):
            clash.append((g, v, prev))
        last[(p, v)] = g
    assert clash and all(g - prev == 4 for g, v, prev in clash)
