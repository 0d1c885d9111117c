/* sparse_matrix.h -- umbrella header (reference include/sparse_matrix.h:15-25). */
#pragma once
#include "CRS.h"
#include "HCSR.h"
#include "SELL_C_Sigma.h"
#include "CSR5.h"
#include "LAV.h"
#include "VNEC.h"
#include "VHCC.h"
#include "CVR.h"
#include "convert.h"
#include "spmv_mtx.h"
