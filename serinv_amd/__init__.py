"""serinv_amd: MI355X-native SPD BTA factorization, selected inversion and solve."""
