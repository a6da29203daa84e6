/* CPU ORACLE -- test infrastructure only (DEM path); filled in below. */
int odem_placeholder(void) { return 0; }
