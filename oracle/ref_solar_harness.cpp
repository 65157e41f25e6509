/*
 * ref_solar_harness.cpp -- a C entry point around the reference's solar
 * position routine (test infrastructure only: it builds a checker for the
 * library's solar_position, nothing in the library links it).
 *
 * The reference's sunAngle (src/implementations/thermal_renderer/gpl/
 * solar_position.cpp:50-90) is compiled where it lies by the Makefile's
 * `ref` target (with -DNDEBUG: its azimuth assertion is not part of the
 * returned value) and linked here; nothing of it is copied.
 */
#include <ctime>
#include <cstring>

#include "solar_position.h"

extern "C" double ref_sun_zenith(double latitude, double longitude, int utc, int year, int month,
                                 int day, int hour, int minute)
{
    struct tm t;
    std::memset(&t, 0, sizeof t);
    t.tm_year = year - 1900;
    t.tm_mon = month - 1;
    t.tm_mday = day;
    t.tm_hour = hour;
    t.tm_min = minute;
    return sunAngle(latitude, longitude, utc, t);
}
