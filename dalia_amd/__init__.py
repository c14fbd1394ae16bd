"""dalia_amd: MI355X-native inner solve of DALIA's f(theta) on the BTA precision."""
