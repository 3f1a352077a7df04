// Time scales shared by the host-side timestamp code (precession.cpp, uvfits.cpp).
#pragma once

namespace sm {
// TAI - UTC [s] at a GPS time, from the IERS leap-second table.
int tai_minus_utc_at(double gps_seconds);
// Julian date on the UTC scale of a GPS time.
double utc_jd_of_gps(double gps_seconds);
}  // namespace sm
