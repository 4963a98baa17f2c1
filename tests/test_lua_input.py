from hande_amd.lua_input import LuaCall, parse


def test_parse_deck():
    env, calls = parse('''
    sys = hubbard_k { electrons = 4, lattice = { {4} }, ms = 0, sym = 'tot_sym', }
    -- comment
    fciqmc { sys = sys, qmc = { tau = 0.01, rng_seed = -7, initiator = true, target_population = 1e5 } }
    --[[ block
    comment ]]
    ''')
    assert isinstance(env["sys"], LuaCall) and env["sys"].func == "hubbard_k"
    assert env["sys"].args["lattice"] == [[4]]
    call = calls[0]
    assert call.func == "fciqmc"
    q = call.args["qmc"]
    assert q["tau"] == 0.01 and q["rng_seed"] == -7 and q["initiator"] is True and q["target_population"] == 1e5
    assert call.args["sys"] is env["sys"]
