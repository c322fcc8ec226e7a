// Triton-format flood run on the APPFIS API.  Reads a Triton case -- the cfg file of
// key=value lines ('#' starts a comment), its DEM (ESRI ASCII with the six-line header,
// or binary with a six-double header), optional initial depth / momentum and Manning
// maps (ASCII, or binary without header), point-source locations ("x,y" lines, '%'
// comment lines) and the hydrograph (comma-separated rows: time in hours, then one flow
// per source in m^3/s, '%' starts a comment) -- and runs the hand-written flood kernels
// (APPFIS::kernels::Flood) while simtime < sim_duration, the loop of the reference
// Triton app (Triton/src/triton_appfis.hpp:316-386).  Prints the step count and, with
// -output, writes output/h.txt, u.txt (qx) and v.txt (qy) with six decimals, as the
// reference app does.  The inputs are mapped onto the grids as the reference app maps
// them (triton_appfis.hpp:99-290): DEM ghost ring 1e6, n squared with its ring copied
// from the inner layer, a source at x, y on full-grid column ceil((x - xllcorner) /
// cellsize) - 1 and row ceil(nrows - (y - yllcorner) / cellsize) - 1 (a later source on
// the same cell replaces an earlier one).  With -ensemble=N (a case without point
// sources) it also advances N copies of the case together through
// APPFIS::kernels::FloodEnsemble and checks that every copy ends identical to the single
// run.
//   flood <case.cfg> [-output] [-fp32] [-plain] [-ensemble=N]
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "Appfis/Appfis.hpp"

namespace
{
    [[noreturn]] void Fail(const std::string& msg)
    {
        std::cerr << "flood: " << msg << std::endl;
        exit(EXIT_FAILURE);
    }

    std::string Strip(const std::string& s)
    {
        const size_t a = s.find_first_not_of(" \t\r\n"), b = s.find_last_not_of(" \t\r\n");
        return a == std::string::npos ? std::string() : s.substr(a, b - a + 1);
    }

    std::string Unquote(const std::string& s)
    {
        if (s.size() >= 2 && s.front() == '"' && s.back() == '"') return s.substr(1, s.size() - 2);
        return s;
    }

    std::map<std::string, std::string> ReadCfg(const std::string& path)
    {
        std::ifstream in(path.c_str());
        if (!in.good()) Fail("cannot read " + path);
        std::map<std::string, std::string> kv;
        std::string line;
        while (std::getline(in, line)) {
            line = Strip(line);
            if (line.empty() || line[0] == '#') continue;
            const size_t hash = line.find('#');
            if (hash != std::string::npos) line = Strip(line.substr(0, hash));
            const size_t eq = line.find('=');
            if (eq == std::string::npos) continue;
            kv[Strip(line.substr(0, eq))] = Unquote(Strip(line.substr(eq + 1)));
        }
        return kv;
    }

    struct Cfg {
        std::map<std::string, std::string> kv;
        std::string Str(const char* k) const
        {
            auto it = kv.find(k);
            return it == kv.end() ? std::string() : it->second;
        }
        double Num(const char* k, double dflt = 0.0) const
        {
            const std::string v = Str(k);
            return v.empty() ? dflt : atof(v.c_str());
        }
    };

    // the six header values: ncols, nrows, xllcorner, yllcorner, cellsize, nodata
    std::vector<double> DemHeader(const std::string& path, bool binary)
    {
        std::vector<double> h(6, 0.0);
        if (binary) {
            std::ifstream in(path.c_str(), std::ios::binary);
            if (!in.good()) Fail("cannot read " + path);
            in.read((char*)h.data(), 6 * sizeof(double));
            if (!in) Fail("short DEM header in " + path);
            return h;
        }
        std::ifstream in(path.c_str());
        if (!in.good()) Fail("cannot read " + path);
        std::string line;
        for (int k = 0; k < 6; k++) {
            if (!std::getline(in, line)) Fail("short DEM header in " + path);
            std::istringstream ss(line);
            std::string tok, last;
            while (ss >> tok) last = tok;
            h[k] = atof(last.c_str());
        }
        return h;
    }

    std::vector<std::vector<double>> ReadRows(const std::string& path, char comment)
    {
        std::ifstream in(path.c_str());
        if (!in.good()) Fail("cannot read " + path);
        std::vector<std::vector<double>> rows;
        std::string line;
        while (std::getline(in, line)) {
            const size_t c = line.find(comment);
            if (c != std::string::npos) line = line.substr(0, c);
            line = Strip(line);
            if (line.empty()) continue;
            std::vector<double> row;
            std::stringstream ss(line);
            std::string tok;
            while (std::getline(ss, tok, ',')) row.push_back(atof(Unquote(Strip(tok)).c_str()));
            rows.push_back(row);
        }
        return rows;
    }

    // Triton's binary files hold doubles (header included), whatever the run's precision:
    // an FP64 run reads them straight into the grid (Array::LoadBinaryFile reads elements
    // of the grid's own type), an FP32 run reads doubles and rounds them.
    template <typename T>
    void LoadInto(APPFIS::Grid<T>& g, const std::string& path, bool binary, int header)
    {
        APPFIS::Array<T>* full = g.GetFullGrid();
        if (!binary) {
            full->LoadAsciiFile(path, header);
        } else if (sizeof(T) == sizeof(double)) {
            full->LoadBinaryFile(path, header);
        } else {
            std::ifstream in(path.c_str(), std::ios::binary);
            if (!in.good()) Fail("cannot read " + path);
            std::vector<double> v((size_t)full->GetSize());
            in.seekg((std::streamoff)header * (std::streamoff)sizeof(double), std::ios::beg);
            in.read((char*)v.data(), (std::streamsize)(v.size() * sizeof(double)));
            if (!in) Fail("short binary file " + path);
            std::copy(v.begin(), v.end(), full->GetData());  // rounds each value to T
        }
    }

    template <typename T>
    bool SameFullGrid(APPFIS::Grid<T>& a, APPFIS::Grid<T>& b)
    {
        APPFIS::Array<T>* x = a.GetFullGrid();
        APPFIS::Array<T>* y = b.GetFullGrid();
        return x->GetSize() == y->GetSize() &&
               std::memcmp(x->GetData(), y->GetData(), sizeof(T) * (size_t)x->GetSize()) == 0;
    }

    template <typename T>
    int Run(const Cfg& cfg, bool output, int variant, int ensemble)
    {
        const bool binary = cfg.Str("input_format") == "BIN";
        const std::string dem_path = cfg.Str("dem_filename");
        const std::vector<double> hdr = DemHeader(dem_path, binary);
        const int nx = (int)hdr[0], ny = (int)hdr[1];
        const double xll = hdr[2], yll = hdr[3], cs = hdr[4];
        if (nx < 1 || ny < 1 || !(cs > 0.0)) Fail("bad DEM header in " + dem_path);

        APPFIS::Grid<T> dem(APPFIS::DIM_2D, nx, ny, APPFIS::SINGLE), nin(APPFIS::DIM_2D, nx, ny, APPFIS::SINGLE);
        APPFIS::Grid<T> hin(APPFIS::DIM_2D, nx, ny, APPFIS::SINGLE), uin(APPFIS::DIM_2D, nx, ny, APPFIS::SINGLE);
        APPFIS::Grid<T> vin(APPFIS::DIM_2D, nx, ny, APPFIS::SINGLE);
        for (APPFIS::Grid<T>* g : {&dem, &nin, &hin, &uin, &vin}) g->AllocateFullGrid();
        LoadInto(dem, dem_path, binary, 6);
        if (!cfg.Str("n_infile").empty()) LoadInto(nin, cfg.Str("n_infile"), binary, 0);
        else nin.GetFullGrid()->FillData((T)cfg.Num("const_mann"));
        nin.GetFullGrid()->Square();
        if (!cfg.Str("h_infile").empty()) LoadInto(hin, cfg.Str("h_infile"), binary, 0);
        if (!cfg.Str("qx_infile").empty()) LoadInto(uin, cfg.Str("qx_infile"), binary, 0);
        if (!cfg.Str("qy_infile").empty()) LoadInto(vin, cfg.Str("qy_infile"), binary, 0);
        for (APPFIS::Grid<T>* g : {&dem, &nin, &hin, &uin, &vin}) APPFIS::Scatter(g->GetFullGrid(), g->GetSubGrid());
        dem.SetBoundary((T)1e6);
        nin.SetBoundaryCopy();

        // point sources: full-grid cell of each location; later sources replace earlier
        std::vector<int> cells;
        std::vector<double> times, flows;
        const int nsrc_cfg = (int)cfg.Num("num_sources");
        if (nsrc_cfg > 0) {
            const std::vector<std::vector<double>> loc = ReadRows(cfg.Str("src_loc_file"), '%');
            const std::vector<std::vector<double>> hyd = ReadRows(cfg.Str("hydrograph_filename"), '%');
            if ((int)loc.size() < nsrc_cfg) Fail("fewer source locations than num_sources");
            if (hyd.empty()) Fail("empty hydrograph");
            std::map<int, int> at;  // full-grid index -> source number (1-based)
            for (int k = 0; k < nsrc_cfg; k++) {
                if (loc[k].size() < 2) Fail("source location needs x,y");
                const int col = (int)std::ceil((loc[k][0] - xll) / cs) - 1;
                const int row = (int)std::ceil(ny - (loc[k][1] - yll) / cs) - 1;
                if (col < 0 || col >= nx || row < 0 || row >= ny) Fail("source outside the DEM");
                at[row * nx + col] = k + 1;
            }
            for (const auto& e : at) {
                const int row = e.first / nx, col = e.first % nx;
                cells.push_back((col + 1) + (nx + 2) * (row + 1));
            }
            for (const auto& r : hyd) {
                times.push_back(r.at(0) * 3600.0);
                for (const auto& e : at) {
                    if ((int)r.size() <= e.second) Fail("hydrograph row without a flow for every source");
                    flows.push_back(r[e.second]);
                }
            }
        }
        appfis_flood_sources src{(int)cells.size(), cells.data(), (int)times.size(), times.data(), flows.data()};
        if (ensemble > 1 && !cells.empty()) Fail("-ensemble needs a case without point sources");

        // -ensemble=N: N copies of the initial state (the DEM and Manning grids are shared)
        std::vector<std::unique_ptr<APPFIS::Grid<T>>> copies;
        for (int k = 0; ensemble > 1 && k < ensemble; k++)
            for (APPFIS::Grid<T>* g : {&hin, &uin, &vin}) {
                copies.emplace_back(new APPFIS::Grid<T>(APPFIS::DIM_2D, nx, ny, APPFIS::SINGLE));
                copies.back()->AllocateFullGrid();
                APPFIS::Array<T>* from = g->GetFullGrid();
                std::copy(from->GetData(), from->GetData() + from->GetSize(), copies.back()->GetFullGrid()->GetData());
                APPFIS::Scatter(copies.back()->GetFullGrid(), copies.back()->GetSubGrid());
            }

        appfis_flood_params prm{};
        prm.dx = cs;
        prm.courant = cfg.Num("courant");
        prm.hextra = cfg.Num("hextra");
        prm.dt_fallback = cfg.Num("time_step");
        prm.print_interval = cfg.Num("print_interval");
        prm.rain_rate = 0.0;
        prm.start_time = cfg.Num("sim_start_time");
        prm.time_increment_fixed = (int)cfg.Num("time_increment_fixed");
        prm.end_time = cfg.Num("sim_duration");
        if (!(prm.end_time > prm.start_time)) Fail("sim_duration must exceed sim_start_time");

        APPFIS::InitializeExecution(dem, nin, hin, uin, vin);
        for (auto& g : copies) APPFIS::InitializeExecution(*g);
        APPFIS::Timer timer;
        timer.Start("COMPUTE");
        const int steps = APPFIS::kernels::Flood(hin, uin, vin, dem, nin, prm, cells.empty() ? nullptr : &src,
                                                 1 << 30, variant);
        timer.Stop("COMPUTE");
        if (!copies.empty()) {
            std::vector<APPFIS::kernels::FloodScenario<T>> scen;
            for (size_t k = 0; k < copies.size(); k += 3)
                scen.push_back({copies[k].get(), copies[k + 1].get(), copies[k + 2].get(), &dem, &nin});
            APPFIS::kernels::FloodEnsemble(scen, prm, 1 << 30);
        }
        APPFIS::FinalizeExecution(hin, uin, vin);
        for (auto& g : copies) APPFIS::FinalizeExecution(*g);
        for (APPFIS::Grid<T>* g : {&hin, &uin, &vin}) APPFIS::Gather(g->GetFullGrid(), g->GetSubGrid());
        if (!copies.empty()) {
            bool same = true;
            for (size_t k = 0; k < copies.size(); k++) {
                APPFIS::Gather(copies[k]->GetFullGrid(), copies[k]->GetSubGrid());
                APPFIS::Grid<T>* single[3] = {&hin, &uin, &vin};
                same = same && SameFullGrid(*copies[k], *single[k % 3]);
            }
            std::cout << "Ensemble of " << ensemble << " copies: " << (same ? "identical" : "DIFFERENT")
                      << " to the single run" << std::endl;
            if (!same) return 1;
        }
        std::cout << "Number of iteration: " << steps << std::endl;
        std::cout << "Compute: " << timer.GetCustomTime("COMPUTE") << std::endl;
        std::cout << "BLUPS: " << timer.CalculateBLUPS("COMPUTE", nx * ny, steps) << std::endl;
        if (output) {
            hin.GetFullGrid()->SaveAsciiFile("./output/h.txt");
            uin.GetFullGrid()->SaveAsciiFile("./output/u.txt");
            vin.GetFullGrid()->SaveAsciiFile("./output/v.txt");
        }
        return 0;
    }
}  // namespace

int main(int argc, char* argv[])
{
    if (argc < 2) Fail("usage: flood <case.cfg> [-output] [-fp32] [-plain] [-ensemble=N]");
    Cfg cfg;
    cfg.kv = ReadCfg(argv[1]);
    const bool output = APPFIS::CheckArgument(argc, argv, "output");
    const bool fp32 = APPFIS::CheckArgument(argc, argv, "fp32");
    const int variant = APPFIS::CheckArgument(argc, argv, "plain") ? APPFIS_PLAIN : APPFIS_CHOSEN;
    const int ensemble = APPFIS::CheckArgument(argc, argv, "ensemble") ? APPFIS::GetArgument(argc, argv, "ensemble") : 0;
    APPFIS::Attribute at;
    at.ghostLayer = 1;
    at.periodic = false;
    APPFIS::Initialize(argc, argv, APPFIS::DIM_2D, APPFIS::DIM_2D, at);
    const int rc = fp32 ? Run<float>(cfg, output, variant, ensemble) : Run<double>(cfg, output, variant, ensemble);
    APPFIS::Finalize();
    return rc;
}
