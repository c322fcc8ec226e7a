/*
 * Attribute.hpp -- run-time attributes of an APPFIS program
 * (fields and defaults of the reference Appfis/Attribute.hpp:16-24).
 */
#ifndef APPFIS_ATTRIBUTE_HPP
#define APPFIS_ATTRIBUTE_HPP

namespace APPFIS
{
    struct Attribute {
        Attribute() : ghostLayer(1), threads(1), periodic(false), overlap(false) {}
        int ghostLayer;  // halo width of every decomposed/periodic face
        int threads;     // host threads (kept for source compatibility)
        bool periodic;   // periodic boundaries on every axis
        bool overlap;    // communication/computation overlap (single rank: no effect)
    };

    inline Attribute ATTRIBUTE;
}  // namespace APPFIS

#endif
