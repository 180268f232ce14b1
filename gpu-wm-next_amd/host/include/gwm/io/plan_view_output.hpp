// Reference include path include/gwm/io/plan_view_output.hpp, served by the host mirror.
#pragma once
#include "gwm/gwm_io.hpp"
