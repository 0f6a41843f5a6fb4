// helm.hpp -- complex 5-point solver (helm.hip): library-level release hook.
#pragma once

namespace cathode {

// Frees the complex solver and its staging buffers (cathode_finalize).
void helm_release();
// frees the microwave power-profile buffers (mw.hip)
void mw_release();
// frees the coupled microwave solve's workspace (mwsolve.hip)
void mws_release();

}  // namespace cathode
