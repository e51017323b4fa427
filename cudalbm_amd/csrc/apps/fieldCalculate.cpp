// fieldCalculate: NaN scan, spatial means, vorticity and divergence of .LBMBin results (CPU only).
int run_field_tool(bool calculate, int argc, char **argv);
int main(int argc, char **argv) { return run_field_tool(true, argc, argv); }
