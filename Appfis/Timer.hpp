/*
 * Timer.hpp -- named wall-clock timers and the GLUP/s metric of the APPFIS API
 * (reference Appfis/Timer.hpp:109-185, 323-403: seconds, accumulated per category,
 * CalculateBLUPS = size * iter / (t * 1e9)).
 *
 * Kernels are queued asynchronously on the device stream, so Stop() first waits for
 * the stream of the running program: a COMPUTE timer then measures the device work it
 * brackets, as the reference's per-iteration synchronisation did implicitly.
 */
#ifndef APPFIS_TIMER_HPP
#define APPFIS_TIMER_HPP

#include <sys/time.h>
#include <unistd.h>

#include <cmath>
#include <ctime>
#include <map>
#include <string>
#include <vector>

#include "Common.hpp"  // detail::SyncDevice: waits for the runtime's stream (no-op before Initialize)

namespace APPFIS
{
    class Timer
    {
    public:
        Timer() { Init(); }

        void Start(std::string category) { gettimeofday(&_allTimers[GetCatIndex(category)], nullptr); }

        void Stop(std::string category)
        {
            detail::SyncDevice();
            timeval t2;
            gettimeofday(&t2, nullptr);
            const int idx = GetCatIndex(category);
            const timeval t1 = _allTimers[idx];
            const unsigned long long us = ((unsigned long long)t2.tv_sec * 1000000ULL + t2.tv_usec)
                                          - ((unsigned long long)t1.tv_sec * 1000000ULL + t1.tv_usec);
            _times[idx] += us;
        }

        void Reset()
        {
            _allTimers.clear();
            _times.clear();
            _timeCats.clear();
            Init();
        }

        double GetTotalTime()
        {
            double total = 0.0;
            for (const auto& kv : _timeCats) total += GetCustomTime(kv.first);
            return total;
        }

        double GetCustomTime(std::string category) { return (double)_times[GetCatIndex(category)] * 1e-6; }

        std::string GetCurrentDate()
        {
            char buf[24];
            time_t raw;
            time(&raw);
            strftime(buf, sizeof(buf), "%D", localtime(&raw));
            return std::string(buf);
        }

        std::string GetHostname()
        {
            char host[1024];
            host[1023] = '\0';
            gethostname(host, 1023);
            return std::string(host);
        }

        int AddNewTimer(std::string category)
        {
            auto it = _timeCats.find(category);
            if (it != _timeCats.end()) return it->second;
            const int id = (int)_times.size();
            _timeCats[category] = id;
            _times.push_back(0ULL);
            _allTimers.push_back(timeval{0, 0});
            return id;
        }

        // billion lattice updates per second over `category`
        double CalculateBLUPS(std::string category, int size, int iter)
        {
            const double t = GetCustomTime(category);
            return ((double)size * iter) / (t * std::pow(10, 9));
        }

    private:
        void Init() {}
        int GetCatIndex(const std::string& category) { return AddNewTimer(category); }

        std::map<std::string, int> _timeCats;
        std::vector<unsigned long long> _times;  // microseconds
        std::vector<timeval> _allTimers;
    };
}  // namespace APPFIS

#endif
