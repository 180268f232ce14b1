// Reference include path include/gwm/state/face_field.hpp, served by the host mirror.
#pragma once
#include "gwm/gwm_dycore.hpp"
