// Frame-output writers in the reference's .pos / .json formats (io.cpp).
#pragma once

namespace sphio {

// fUtility.h:254-303 / 459-484; kind is "fluid" or "solid".  Arrays are
// n records (3 floats per vector); unused components may be null.
int write_pos(const char* path, const char* logName, long long frame, const char* kind, int index,
              long long n, unsigned comps, const float* pos, const float* vel, const float* rho,
              const float* prs, const float* acc);
// fUtility.h:390-427
int write_json(const char* path, long long n, const float* pos, const float* vel, float d);

}  // namespace sphio
