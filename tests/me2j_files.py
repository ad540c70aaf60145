"""Synthetic me2j interaction files for the tests (the reference ships none:
its chi2b_srg0625_eMax08_hwHO020.me2j.gz is listed as a missing blob).  The
file layout and the synthetic values live in the package
(eom_imsrg.interaction.write_me2j / me2j_quadruples, eom_imsrg.synth.me2j_values);
these names keep the tests' vocabulary."""
from eom_imsrg.interaction import me2j_quadruples as quadruples, write_me2j as _write
from eom_imsrg.synth import me2j_values as synthetic_values


def write_me2j(path, values, compress=True):
    return _write(path, values, compress=compress, header="synthetic me2j interaction file")


__all__ = ["quadruples", "synthetic_values", "write_me2j"]
