"""Recipe for the generated cell-model headers under chaste_amd/csrc/models.

Each entry names the reference's CellML file (heart/src/odes/cellml/) and the
constants exposed as run-time parameters, in the order of the header's
parameter planes.  Models in seconds or with absolute currents are converted
by the generator (rates per ms; the stimulus through the model's membrane
capacitance).  Regenerate every header with

    python -m chaste_amd.codegen.models [<cellml directory>]

(default: the reference tree's heart/src/odes/cellml under /root/reference).
The O'Hara-Rudy header (ord_endo.h) is written by hand: the reference holds
no CellML for it.
"""
import os
import sys

from chaste_amd.codegen import cellml2cpp

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(os.path.dirname(HERE), 'csrc', 'models')
DEFAULT_CELLML = '/root/reference/heart/src/odes/cellml'

MODELS = {
    'lr91': ('LuoRudy1991.cellml',
             ['fast_sodium_current.g_Na', 'time_dependent_potassium_current.g_Kmax', 'slow_inward_current.P_si']),
    'hh52': ('HodgkinHuxley1952.cellml',
             ['sodium_channel.g_Na', 'potassium_channel.g_K', 'leakage_current.g_L']),
    'fox2002': ('FoxModel2002.cellml',
                ['fast_sodium_current.g_Na', 'rapid_activating_delayed_rectifiyer_K_current.g_Kr',
                 'slow_activating_delayed_rectifiyer_K_current.g_Ks', 'time_independent_potassium_current.g_K1',
                 'L_type_Ca_current.P_Ca']),
    'mahajan2008': ('Mahajan2008.cellml',
                    ['INa.gna', 'ICaL.gca', 'IKr.ScaleFactorGkr', 'IKs.ScaleFactorGks', 'Ito.ScaleFactorIto']),
    'dn85': ('DiFrancescoNoble1985.cellml',
             ['fast_sodium_current.g_Na', 'time_independent_potassium_current.g_K1', 'second_inward_current.P_si',
              'transient_outward_current.g_to']),
    'fr2000': ('FaberRudy2000.cellml',
               ['slow_delayed_rectifier_potassium_current.ScaleFactorGks', 'transient_outward_current.ScaleFactorIto',
                'rapid_delayed_rectifier_potassium_current.ScaleFactorGkr', 'fast_sodium_current.g_Na']),
    'maleckar2008': ('Maleckar2008.cellml',
                     ['sodium_current.ScaleFactorGna', 'L_type_Ca_channel.ScaleFactorGCaL',
                      'delayed_rectifier_K_currents.ScaleFactorGkr', 'delayed_rectifier_K_currents.ScaleFactorGks',
                      'Ca_independent_transient_outward_K_current.ScaleFactorIto',
                      'ultra_rapid_K_current.ScaleFactorGKur']),
    'n98': ('NobleVargheseKohlNoble1998a.cellml',
            ['fast_sodium_current.g_Na', 'time_independent_potassium_current.g_K1', 'L_type_Ca_channel.P_Ca_L',
             'slow_delayed_rectifier_potassium_current.g_Ks', 'rapid_delayed_rectifier_potassium_current.g_Kr1']),
    'tt06_epi': ('TenTusscher2006Epi.cellml',
                 ['fast_sodium_current.g_Na', 'L_type_Ca_current.g_CaL', 'rapid_time_dependent_potassium_current.g_Kr',
                  'slow_time_dependent_potassium_current.g_Ks', 'inward_rectifier_potassium_current.g_K1',
                  'transient_outward_current.g_to']),
}


def generate(cellml_dir=DEFAULT_CELLML, out_dir=OUT, names=None):
    for ns in names or sorted(MODELS):
        fname, params = MODELS[ns]
        argv = [os.path.join(cellml_dir, fname), os.path.join(out_dir, ns + '.h'), ns]
        for p in params:
            argv += ['--param', p]
        cellml2cpp.main(argv)


if __name__ == '__main__':
    generate(*(sys.argv[1:2] or [DEFAULT_CELLML]))
