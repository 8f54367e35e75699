"""cintx_amd: MI355X-native Rys two-electron engine (libcintx_hip.so)."""
