// Reference include path include/gwm/core/dry_thermo.hpp, served by the host mirror.
#pragma once
#include "gwm/gwm_ingest.hpp"
