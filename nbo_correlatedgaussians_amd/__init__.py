"""MI355X-native FP64 energy-and-gradient engine for non-Born-Oppenheimer
explicitly correlated Gaussians (the outprod/matel/gradel/make_grad path of
NECG)."""

from .system import (System, read_input, write_input, format_input,
                     parse_tokens, lih_system, he_system, h2_system,
                     be_system, c_system, hydrogen_atom, generate_parameters, pack_x, unpack_x,
                     lih_oldquit_basis, generate_like_template, format_generated_input,
                     generated_lih_input)

__all__ = ["System", "read_input", "write_input", "format_input",
           "parse_tokens", "lih_system", "he_system", "h2_system",
           "be_system", "c_system", "hydrogen_atom", "generate_parameters", "pack_x", "unpack_x",
           "lih_oldquit_basis", "generate_like_template", "format_generated_input",
           "generated_lih_input"]
