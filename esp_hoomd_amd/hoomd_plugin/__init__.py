"""HOOMD-blue plugin of the MI355X ESP library (needs HOOMD-blue and the _esp module built by
CMakeLists.txt here); import esp_hoomd_amd.hoomd_plugin.esp for Spectral / Local."""
