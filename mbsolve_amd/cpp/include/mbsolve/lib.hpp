/* Umbrella header of the mbsolve API (reference: mbsolve-lib/include/
 * mbsolve/lib.hpp). */
#ifndef MBSOLVE_LIB_H
#define MBSOLVE_LIB_H

#include <mbsolve/lib/device.hpp>
#include <mbsolve/lib/material.hpp>
#include <mbsolve/lib/qm_description.hpp>
#include <mbsolve/lib/record.hpp>
#include <mbsolve/lib/result.hpp>
#include <mbsolve/lib/scenario.hpp>
#include <mbsolve/lib/solver.hpp>
#include <mbsolve/lib/source.hpp>
#include <mbsolve/lib/types.hpp>
#include <mbsolve/lib/writer.hpp>

#endif
