// Reference include path include/gwm/state/tracer_registry.hpp, served by the host mirror.
#pragma once
#include "gwm/gwm_dycore.hpp"
