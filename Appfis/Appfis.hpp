/*
 * Appfis.hpp -- single include of the APPFIS API on MI355X (reference Appfis/Appfis.hpp).
 *
 * Build an application with
 *   hipcc --offload-arch=gfx950 -std=c++17 -DAPPFIS_HIP [-DAPPFIS_CPU] -I<repo>/include ...
 *         -L<repo>/appfis_amd -lappfis_hip -Wl,-rpath,<repo>/appfis_amd
 */
#ifndef APPFIS_HPP
#define APPFIS_HPP

#include "Env.hpp"
#include "Common.hpp"
#include "Grid.hpp"
#include "Runtime.hpp"
#include "Communication.hpp"
#include "Timer.hpp"
#include "Kernels_HIP.hpp"

#endif
