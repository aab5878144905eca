"""femorax_amd: an MI355X-native hex8 matrix-free operator path (see include/femorax_amd.h)."""
