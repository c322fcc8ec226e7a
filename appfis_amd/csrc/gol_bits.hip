// GameOfLife CHOSEN form: bit-packed cells, several generations per launch.
#include "common.hpp"

int gol_bits_run(appfis_ctx* ctx, uint8_t* b0, uint8_t* b1, int nx, int ny, int gens, int tsteps)
{
    (void)ctx; (void)b0; (void)b1; (void)nx; (void)ny; (void)gens; (void)tsteps;
    appfis::set_error("gol: the bit-packed form is not built yet");
    return APPFIS_ERR_UNSUPPORTED;
}
