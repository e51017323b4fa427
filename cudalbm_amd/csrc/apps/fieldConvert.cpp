// fieldConvert: .LBMBin results to VTU / VTS / Tecplot, optionally a cut plane (CPU only).
int run_field_tool(bool calculate, int argc, char **argv);
int main(int argc, char **argv) { return run_field_tool(false, argc, argv); }
